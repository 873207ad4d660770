"""Pinning of the elastic 2D oracle (CPU, no GPU).

The oracle restates elastic2D_iso_xPU.jl, freesurface_derivatives_4th_
mirror.jl, correlate_gradient_xPU.jl (elastic), ela_gradient.jl and the sinc
point spreading of utils.jl.  The reference's own elastic test
(test_gradient_elastic_homogeneous.jl) only checks that gradients are
non-zero and that checkpointing does not change them; here the oracle is
pinned by mathematics:

* seismograms of moment-tensor and point-force sources in a homogeneous
  full space against the closed-form 2D elastodynamic solution (P and S
  potentials through Hankel functions, evaluated per frequency);
* the returned "rho", "lambda", "mu" gradients against central finite
  differences of the discrete L2 misfit (free surface on and off);
* the free surface through the Rayleigh-wave speed of a homogeneous
  half-space (root of the Rayleigh equation);
* Kaiser-windowed sinc taps against their closed forms.
"""
import numpy as np
import pytest
from scipy.optimize import brentq
from scipy.special import hankel2, i0 as scipy_i0


# ---------------------------------------------------------------------------
# Sinc / Kaiser point spreading (utils.jl:71-214)
# ---------------------------------------------------------------------------
def test_kaiser_sinc_closed_forms(oracle_mod):
    o = oracle_mod
    beta, r = 6.31, 4
    assert o.kaiser(0.0, r, beta) == pytest.approx(1.0)
    assert o.kaiser(r, r, beta) == pytest.approx(1.0 / scipy_i0(beta), rel=1e-14)
    assert o.kaiser(1.3, r, beta) == pytest.approx(scipy_i0(beta * np.sqrt(1 - (1.3 / r) ** 2)) / scipy_i0(beta),
                                                   rel=1e-14)
    assert o.kaiser(r + 1e-9, r, beta) == 0.0
    # on a grid point the band-limited delta is a single unit tap
    idx, c = o.coeffsinc1d(35.0, 5.0, 40, r, beta, 0.0, False, 0.0, 195.0)
    assert idx.tolist() == [8] and c.tolist() == [1.0]
    # off the grid: the 2r taps inside the window, closed form kaiser * sinc, nearly unit sum
    x0 = 101.7
    idx, c = o.coeffsinc1d(x0, 5.0, 40, r, beta, 0.0, False, 0.0, 195.0)
    xs = (idx - 1) * 5.0
    ref = [scipy_i0(beta * np.sqrt(1 - ((x - x0) / (r * 5.0)) ** 2)) / scipy_i0(beta) * np.sinc((x - x0) / 5.0)
           for x in xs]
    assert len(idx) == 2 * r
    np.testing.assert_allclose(c, ref, rtol=1e-13)
    assert abs(c.sum() - 1.0) < 2e-3
    # the staggered shift moves the tap grid by half a cell
    idx_s, c_s = o.coeffsinc1d(x0 + 2.5, 5.0, 40, r, beta, 2.5, False, 0.0, 195.0)
    np.testing.assert_array_equal(idx_s, idx)
    np.testing.assert_allclose(c_s, c, rtol=1e-12)
    # near the boundary taps fold back; `mirror` flips their sign
    i_m, c_m = o.coeffsinc1d(6.0, 5.0, 40, r, beta, 0.0, True, 0.0, 195.0)
    i_p, c_p = o.coeffsinc1d(6.0, 5.0, 40, r, beta, 0.0, False, 0.0, 195.0)
    np.testing.assert_array_equal(i_m, i_p)
    assert i_m.min() >= 1
    unfolded = {k: o.modd((k - 1) * 5.0, 6.0, r, beta, 5.0) for k in range(-5, 8)}
    # tap 3 (x = 10) receives its own weight plus the fold of x = -10 (tap -1)
    k3 = list(i_m).index(3)
    assert c_p[k3] == pytest.approx(unfolded[3] + unfolded[-1], rel=1e-12)
    assert c_m[k3] == pytest.approx(unfolded[3] - unfolded[-1], rel=1e-12)


# ---------------------------------------------------------------------------
# Homogeneous full space: closed-form 2D elastodynamic solution
# ---------------------------------------------------------------------------
def _analytic_fk(ricker, rho, vp, vs, dt, nt, t0, f0, src, rec, M=None, F=None, nfft=4096, d=0.25):
    """u(t) at `rec` (times dt*1..nt) for rho u_tt = div(C:eps + M delta s) + F delta s.

    With u = grad phi + (d_z psi, -d_x psi) and Q_k = ln r / 2pi - (i/4) H0^(2)(k r)
    (so that (Lap + k^2) Q_k = Lap^-1 delta times k^2, e^{+i w t} convention):
      phi = -s/(rho w^2) (Mxx dxx + Mzz dzz + 2 Mxz dxz + Fx dx + Fz dz) Q_kp
      psi = -s/(rho w^2) (Mxz (dzz - dxx) + (Mxx - Mzz) dxz + Fx dz - Fz dx) Q_ks
    Spatial derivatives by central differences of step d (error (k d)^2)."""
    s = ricker(dt * np.arange(nfft), t0, f0)
    S = np.fft.rfft(s) * dt
    w = 2 * np.pi * np.fft.rfftfreq(nfft, dt)[1:]
    x, z = rec[0] - src[0], rec[1] - src[1]

    def Q(dx_, dz_, k):
        r = np.hypot(x + dx_, z + dz_)
        return np.log(r) / (2 * np.pi) - 0.25j * hankel2(0, k * r)

    def D(k, ax, az):
        # central difference operator of order (ax, az) applied to Q at (x, z)
        pts = {(0, 0): 1.0}
        for _ in range(ax):
            nxt = {}
            for (a, b), cf in pts.items():
                nxt[(a + 1, b)] = nxt.get((a + 1, b), 0) + cf / (2 * d)
                nxt[(a - 1, b)] = nxt.get((a - 1, b), 0) - cf / (2 * d)
            pts = nxt
        for _ in range(az):
            nxt = {}
            for (a, b), cf in pts.items():
                nxt[(a, b + 1)] = nxt.get((a, b + 1), 0) + cf / (2 * d)
                nxt[(a, b - 1)] = nxt.get((a, b - 1), 0) - cf / (2 * d)
            pts = nxt
        return sum(cf * Q(a * d, b * d, k) for (a, b), cf in pts.items() if cf != 0)

    kp, ks = w / vp, w / vs
    Mxx, Mzz, Mxz = M if M is not None else (0.0, 0.0, 0.0)
    Fx, Fz = F if F is not None else (0.0, 0.0)

    def phi(ax, az):
        return (Mxx * D(kp, ax + 2, az) + Mzz * D(kp, ax, az + 2) + 2 * Mxz * D(kp, ax + 1, az + 1)
                + Fx * D(kp, ax + 1, az) + Fz * D(kp, ax, az + 1))

    def psi(ax, az):
        return (Mxz * (D(ks, ax, az + 2) - D(ks, ax + 2, az)) + (Mxx - Mzz) * D(ks, ax + 1, az + 1)
                + Fx * D(ks, ax, az + 1) - Fz * D(ks, ax + 1, az))

    c = -S[1:] / (rho * w ** 2)
    U = np.zeros((len(w) + 1, 2), complex)
    U[1:, 0] = c * (phi(1, 0) + psi(0, 1))
    U[1:, 1] = c * (phi(0, 1) - psi(1, 0))
    u = np.fft.irfft(U, nfft, axis=0) / dt
    return u[1:nt + 1]


@pytest.fixture(scope="module")
def homogeneous():
    rho, vp = 2000.0, 3000.0
    vs = vp / np.sqrt(3)
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    n, dx = 301, 5.0
    dt = 0.5 * dx / vp / (7 / 6) / np.sqrt(2)
    return dict(rho=rho, vp=vp, vs=vs, lam=lam, mu=mu, n=n, dx=dx, dt=dt, nt=520, f0=15.0, t0=0.1)


@pytest.mark.parametrize("source", ["explosion", "shear", "general", "force"])
def test_el_forward_matches_closed_form(oracle_mod, homogeneous, source):
    o = oracle_mod
    p = homogeneous
    n, dx, dt, nt = p["n"], p["dx"], p["dt"], p["nt"]
    lx = (n - 1) * dx
    src = np.array([lx / 2, lx / 2])
    recs = np.array([[lx / 2 + 350.0, lx / 2 + 12.0], [lx / 2 - 140.0, lx / 2 + 310.0]])
    tf = o.rickerstf(dt * np.arange(nt), p["t0"], p["f0"])
    M, F = None, None
    if source == "force":
        F = (1e6, -4e5)
        stf = np.zeros((nt, 2, 1))
        stf[:, 0, 0] = F[0] * tf
        stf[:, 1, 0] = F[1] * tf
        shot = dict(src_pos=src[None], srctf=stf, f0=p["f0"], rec_pos=recs)
    else:
        M = {"explosion": (1e6, 1e6, 0.0), "shear": (0.0, 0.0, 1e6), "general": (1e6, -5e5, 3e5)}[source]
        shot = dict(src_pos=src[None], srctf=tf[:, None], moments=np.array([M]), f0=p["f0"], rec_pos=recs)
    full = np.ones((n, n))
    seis = o.elastic_forward(p["lam"] * full, p["mu"] * full, p["rho"] * full, (dx, dx), dt, nt, 20, 1e-4, False,
                             [shot])[0]
    for r in range(recs.shape[0]):
        ref = _analytic_fk(o.rickerstf, p["rho"], p["vp"], p["vs"], dt, nt, p["t0"], p["f0"], src, recs[r], M=M, F=F)
        err = np.linalg.norm(seis[:, :, r] - ref) / np.linalg.norm(ref)
        print(source, r, err)
        assert err < 1e-2


# ---------------------------------------------------------------------------
# Gradient: central finite differences of the discrete misfit
# ---------------------------------------------------------------------------
def _small_el_problem(o, freetop, kind, seed=0):
    rng = np.random.default_rng(seed)
    nx, nz, h = 56, 48, 10.0
    dt, nt, f0, halo = 1e-3, 260, 12.0, 8

    def smooth(a):
        from scipy.ndimage import gaussian_filter
        return gaussian_filter(a, 2.0)

    rho = smooth(2000 + 200 * rng.random((nx, nz)))
    vp = smooth(2800 + 300 * rng.random((nx, nz)))
    vs = vp / np.sqrt(3) * smooth(0.9 + 0.1 * rng.random((nx, nz)))
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    lam0, mu0, rho0 = [np.full_like(a, a.mean()) for a in (lam, mu, rho)]
    # a corner cell fixes max(vp), which sets the C-PML profile
    vpm = float(np.sqrt((lam + 2 * mu) / rho).max()) * 1.05
    for L, Mu, R in ((lam, mu, rho), (lam0, mu0, rho0)):
        R[1, 1] = 2000.0
        Mu[1, 1] = 2000.0 * (vpm / np.sqrt(3)) ** 2
        L[1, 1] = 2000.0 * vpm ** 2 - 2 * Mu[1, 1]
    tf = o.rickerstf(dt * np.arange(nt), 1.2 / f0, f0)
    src = np.array([[25 * h, 20 * h]])
    rec = np.array([[x * h + 3.0, 12 * h] for x in range(12, 46, 5)] + [[30 * h, 35 * h + 4.0]])
    if kind == "ef":
        stf = np.zeros((nt, 2, 1))
        stf[:, 0, 0] = tf
        stf[:, 1, 0] = 0.7 * tf
        shot = dict(src_pos=src, srctf=stf, f0=f0, rec_pos=rec)
    else:
        shot = dict(src_pos=src, srctf=tf[:, None], moments=np.array([[1.0, 0.5, 0.3]]), f0=f0, rec_pos=rec)
    args = ((h, h), dt, nt, halo, 1e-4, freetop)
    shot["observed"] = o.elastic_forward(lam, mu, rho, *args, [shot])[0]
    return dict(lam=lam0, mu=mu0, rho=rho0, args=args, shot=shot)


@pytest.mark.parametrize("freetop,kind", [(False, "ef"), (True, "mt"), (True, "ef")])
def test_el_gradient_matches_central_fd(oracle_mod, freetop, kind):
    """The reference's elastic gradients are dJ/dm of the discrete misfit
    (note: + sign, unlike the acoustic drivers which return -dJ/dm)."""
    o = oracle_mod
    p = _small_el_problem(o, freetop, kind)
    g, _, _ = o.elastic_gradient(p["lam"], p["mu"], p["rho"], *p["args"], [p["shot"]], check_freq=20)

    def J(m):
        syn = o.elastic_forward(m["lam"], m["mu"], m["rho"], *p["args"], [p["shot"]])[0]
        return o.l2_misfit(syn, p["shot"]["observed"])

    for key, base in (("lambda", "lam"), ("mu", "mu"), ("rho", "rho")):
        for (i, j) in [(20, 25), (40, 30), (30, 15)]:
            m = {k: p[k].copy() for k in ("lam", "mu", "rho")}
            d = 1e-4 * abs(p[base][i, j])
            m[base][i, j] += d
            jp = J(m)
            m[base][i, j] -= 2 * d
            jm = J(m)
            fd = (jp - jm) / (2 * d)
            assert abs(fd) > 0
            assert g[key][i, j] == pytest.approx(fd, rel=1e-6), (key, i, j)


@pytest.mark.parametrize("freetop", [False, True])
def test_el_checkpointing_is_exact(oracle_mod, freetop):
    o = oracle_mod
    p = _small_el_problem(o, freetop, "mt")
    outs = [o.elastic_gradient(p["lam"], p["mu"], p["rho"], *p["args"], [p["shot"]], check_freq=cf)
            for cf in (1, 7, 50)]
    for g, mis, seis in outs[1:]:
        for k in ("rho", "lambda", "mu"):
            assert np.array_equal(g[k], outs[0][0][k])
        assert mis == outs[0][1]
        assert np.array_equal(seis[0], outs[0][2][0])
    # the gradient path's forward is swforward's
    fwd = o.elastic_forward(p["lam"], p["mu"], p["rho"], *p["args"], [p["shot"]])[0]
    assert np.array_equal(fwd, outs[0][2][0])
    assert all(np.any(outs[0][0][k] != 0) for k in ("rho", "lambda", "mu"))


def test_el_vector_misfit_closed_form(oracle_mod):
    o = oracle_mod
    rng = np.random.default_rng(1)
    syn, obs = rng.standard_normal((50, 2, 3)), rng.standard_normal((50, 2, 3))
    r = syn - obs
    assert o.l2_misfit(syn, obs) == pytest.approx(0.5 * np.sum(r * r), rel=1e-14)
    C = np.diag(rng.random(50) + 0.5)
    want = 0.5 * sum(np.trace(r[:, c, :].T @ C @ r[:, c, :]) for c in range(2))
    assert o.l2_misfit(syn, obs, C) == pytest.approx(want, rel=1e-13)
    np.testing.assert_allclose(o.l2_dchi_du(syn, obs, C)[:, 1, :], C @ r[:, 1, :], rtol=1e-14)


# ---------------------------------------------------------------------------
# Free surface: Rayleigh wave speed on a homogeneous half-space
# ---------------------------------------------------------------------------
def test_el_free_surface_rayleigh_speed(oracle_mod):
    o = oracle_mod
    rho, vp = 2000.0, 3000.0
    vs = vp / np.sqrt(3)
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    nx, nz, dx = 600, 200, 5.0
    dt = 0.5 * dx / vp / (7 / 6) / np.sqrt(2)
    f0, t0 = 15.0, 0.1
    nt = int(1.1 / dt)
    tf = o.rickerstf(dt * np.arange(nt), t0, f0)
    stf = np.zeros((nt, 2, 1))
    stf[:, 1, 0] = tf
    xs = 500.0
    rec = np.array([[xs + 400.0, 0.0], [xs + 1200.0, 0.0], [xs + 1200.0, 100.0]])
    shot = dict(src_pos=np.array([[xs, 0.0]]), srctf=stf, f0=f0, rec_pos=rec)
    full = np.ones((nx, nz))
    seis = o.elastic_forward(lam * full, mu * full, rho * full, (dx, dx), dt, nt, 20, 1e-4, True, [shot])[0]
    a, b = seis[:, 1, 0], seis[:, 1, 1]
    cc = np.correlate(b, a, "full")
    k = int(np.argmax(cc))
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    delay = (k - (nt - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)) * dt
    q = (vs / vp) ** 2
    xr = brentq(lambda x: x ** 6 - 8 * x ** 4 + (24 - 16 * q) * x ** 2 - 16 * (1 - q), 0.5, 0.99)
    assert 800.0 / delay == pytest.approx(xr * vs, rel=1e-2)
    # surface wave: larger at the surface than 100 m below
    assert np.abs(seis[:, 1, 1]).max() > 3 * np.abs(seis[:, 1, 2]).max()


# ---------------------------------------------------------------------------
# Free surface: exact half-space response of a buried explosive line source
# ---------------------------------------------------------------------------
def _eta(k2, xi2):
    """Vertical wavenumber with Im <= 0 (e^{-i eta |z|} decays, e^{+i w t})."""
    e = np.sqrt(k2 - xi2)
    return np.where(e.imag > 0, -e, e)


def _halfspace_explosion(ricker, rho, vp, vs, dt, nt, t0, f0, zs, pts, M0, nfft=4096, nxi=8000):
    """u(t) (times dt*1..nt) at points (x, z), 0 <= z < zs, x measured from
    the source, of an explosive line source (Mxx = Mzz = M0, the oracle's
    moment-tensor injection) at depth zs below the free surface z = 0 of a
    homogeneous half-space, by wavenumber integration (Bouchon's discrete
    wavenumber method: complex frequency w - i eps, eps = pi / T).

    The incident P potential phi = -(M0 S / (rho vp^2)) (i/4) H0^(2)(kp r)
    (the same source as the full-space closed form above) is expanded in
    plane waves, H0^(2)(k r) = (1/pi) int e^{-i xi x - i eta |z - zs|} / eta
    dxi; each upgoing plane wave of amplitude P at z = 0 is reflected as a P
    wave Rp and an SV wave Rs, fixed by sigma_zz = sigma_xz = 0 at z = 0:
      -(lam kp^2 + 2 mu eta_p^2) Rp + 2 mu xi eta_s Rs = (lam kp^2 + 2 mu eta_p^2) P
      -2 xi eta_p Rp + (xi^2 - eta_s^2) Rs = -2 xi eta_p P
    with u = grad phi + (d_z psi, -d_x psi)."""
    t = dt * np.arange(nfft)
    eps = np.pi / (nfft * dt)
    S = np.fft.rfft(ricker(t, t0, f0) * np.exp(-eps * t)) * dt
    wc = 2 * np.pi * np.fft.rfftfreq(nfft, dt) - 1j * eps
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    ximax = 1.6 * 2 * np.pi * (4 * f0) / vs
    xi = np.linspace(-ximax, ximax, nxi)
    dxi = xi[1] - xi[0]
    px = np.array([q[0] for q in pts])
    pz = np.array([q[1] for q in pts])
    ex = np.exp(-1j * np.outer(px, xi))
    U = np.zeros((len(wc), len(pts), 2), complex)
    for iw, w in enumerate(wc):
        kp2, ks2 = (w / vp) ** 2, (w / vs) ** 2
        ep, es = _eta(kp2, xi ** 2), _eta(ks2, xi ** 2)
        P = -(M0 / (rho * vp ** 2)) * 0.25j / np.pi * S[iw] * np.exp(-1j * ep * zs) / ep
        c = lam * kp2 + 2 * mu * ep ** 2
        a11, a12, b1 = -c, 2 * mu * xi * es, c * P
        a21, a22, b2 = -2 * xi * ep, xi ** 2 - es ** 2, -2 * xi * ep * P
        det = a11 * a22 - a12 * a21
        Rp = (b1 * a22 - a12 * b2) / det
        Rs = (a11 * b2 - a21 * b1) / det
        up, us = np.exp(1j * np.outer(pz, ep)), np.exp(-1j * np.outer(pz, es))
        ux = -1j * xi * (P * up + Rp / up) - 1j * es * Rs * us
        uz = 1j * ep * (P * up - Rp / up) + 1j * xi * Rs * us
        U[iw, :, 0] = np.sum(ex * ux, 1) * dxi
        U[iw, :, 1] = np.sum(ex * uz, 1) * dxi
    u = np.fft.irfft(U, nfft, axis=0) / dt * np.exp(eps * t)[:, None, None]
    return u[1:nt + 1]


def test_halfspace_expansion_reduces_to_full_space(oracle_mod, homogeneous):
    """Without the reflected waves the plane-wave expansion is the
    full-space closed form (pins the expansion, its branch and scaling)."""
    p = homogeneous
    zs, pt = 300.0, (200.0, 150.0)
    full = _analytic_fk(oracle_mod.rickerstf, p["rho"], p["vp"], p["vs"], p["dt"], p["nt"], p["t0"], p["f0"],
                        np.array([0.0, zs]), np.array(pt), M=(1e6, 1e6, 0.0))
    # incident field alone: the half-space formula with the reflection
    # coefficients zero, i.e. evaluated far above a source in a full space
    t = p["dt"] * np.arange(4096)
    eps = np.pi / (4096 * p["dt"])
    S = np.fft.rfft(oracle_mod.rickerstf(t, p["t0"], p["f0"]) * np.exp(-eps * t)) * p["dt"]
    wc = 2 * np.pi * np.fft.rfftfreq(4096, p["dt"]) - 1j * eps
    xi = np.linspace(-1.6 * 2 * np.pi * 60 / p["vs"], 1.6 * 2 * np.pi * 60 / p["vs"], 8000)
    U = np.zeros((len(wc), 2), complex)
    for iw, w in enumerate(wc):
        ep = _eta((w / p["vp"]) ** 2, xi ** 2)
        P = -(1e6 / (p["rho"] * p["vp"] ** 2)) * 0.25j / np.pi * S[iw] * np.exp(-1j * ep * zs) / ep
        e = np.exp(-1j * xi * pt[0] + 1j * ep * pt[1]) * (xi[1] - xi[0])
        U[iw] = [np.sum(e * -1j * xi * P), np.sum(e * 1j * ep * P)]
    inc = (np.fft.irfft(U, 4096, axis=0) / p["dt"] * np.exp(eps * t)[:, None])[1:p["nt"] + 1]
    assert np.linalg.norm(inc - full) / np.linalg.norm(full) < 1e-4


def test_el_free_surface_matches_halfspace_solution(oracle_mod):
    """Lamb's problem for a buried explosive line source (the 2D analogue of
    Garvin's problem): the oracle's displacements on and just below the free
    surface against the exact half-space response.  Below the surface
    (25 and 40 m, clear of the receivers' fold at the surface) each trace is
    within rel-L2 4e-3; on the surface within the reference's own 1%
    integrated-error bound (test_analytical_vs_numerical_acoustic_constant_
    density.jl:38; uz there carries the first-order error of the receiver
    taps folded across the surface).  Four copies of the oracle, each with
    ONE free-surface mirror sign of freesurface_derivatives_4th_mirror.jl
    flipped (ORACLE_PLANT in oracle_el2d.h), each miss the 4e-3 bound by at
    least 1.5x on some buried trace."""
    import ctypes
    import os
    import subprocess
    o = oracle_mod
    rho, vp = 2000.0, 3000.0
    vs = vp / np.sqrt(3)
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    f0, t0, M0, dx = 15.0, 0.1, 1e6, 2.5
    dt = 0.5 * dx / vp / (7 / 6) / np.sqrt(2)
    nt = int(0.5 / dt)
    nx, nz = int(1500 / dx) + 1, int(500 / dx) + 1
    xs, zs = 750.0, 100.0
    pts = [(150.0, 0.0), (300.0, 0.0), (150.0, 25.0), (300.0, 25.0), (300.0, 40.0)]
    rec = np.array([[xs + x, z] for x, z in pts])
    tf = o.rickerstf(dt * np.arange(nt), t0, f0)
    shot = dict(src_pos=np.array([[xs, zs]]), srctf=tf[:, None], moments=np.array([[M0, M0, 0.0]]), f0=f0,
                rec_pos=rec)
    full = np.ones((nx, nz))
    ref = _halfspace_explosion(o.rickerstf, rho, vp, vs, dt, nt, t0, f0, zs, pts, M0)

    def errors():
        seis = o.elastic_forward(lam * full, mu * full, rho * full, (dx, dx), dt, nt, 20, 1e-4, True, [shot])[0]
        l2 = [[np.linalg.norm(seis[:, c, r] - ref[:, r, c]) / np.linalg.norm(ref[:, r, c]) for c in range(2)]
              for r in range(len(pts))]
        integ = [[np.sum(np.abs(seis[:, c, r] - ref[:, r, c])) / (np.abs(ref[:, r, c]).max() * nt)
                  for c in range(2)] for r in range(len(pts))]
        return np.array(l2), np.array(integ)

    l2, integ = errors()
    print("free surface, buried receivers rel-L2:", l2[2:].ravel(), "surface integrated:", integ[:2].ravel())
    assert l2[2:].max() <= 4e-3
    assert integ[:2].max() <= 1e-2
    subprocess.run(["make", "-s", "-C", os.path.dirname(o.__file__), "planted"], check=True)
    good = o.lib()
    try:
        for k in (1, 2, 3, 4):
            o._LIB = ctypes.CDLL(os.path.join(os.path.dirname(o.__file__), "_build", f"liboracle_plant{k}.so"))
            l2k, _ = errors()
            print(f"mirror sign {k} flipped: buried rel-L2 max {l2k[2:].max():.2e}")
            assert l2k[2:].max() >= 1.5 * 4e-3
    finally:
        o._LIB = good
