"""Pins the variable-density (staggered) part of the CPU oracle: the
reference's own analytic 2D test and tolerance, checkpoint exactness, and
central finite differences of the misfit for the vp and rho gradients."""
import numpy as np
import pytest


@pytest.mark.parametrize("halo,rcoef,nt", [(0, 1.0, 350), (20, 1e-4, 1400)])
def test_vd_forward_analytic_2d(oracle_mod, halo, rcoef, nt):
    """test_analytical_vs_numerical_acoustic_variable_density.jl:68-113 with
    setup_models.jl:107-130 and analytical_solution_constant_vel_constant_
    density_2D (setup_models.jl:240-258): Gaussian-derivative source, half
    step time axis, 1% integrated-error bound."""
    o = oracle_mod
    c0, rho0, n, dx, f0 = 1000.0, 1500.0, 401, 5.0, 5.0
    dt = 0.99 * dx / c0 / np.sqrt(2) * 6 / 7
    t0 = 2 / f0
    lx = (n - 1) * dx
    tf = o.gaussderivstf(dt * np.arange(nt), t0, f0)[:, None]
    src, rec = np.array([[lx / 2, lx / 2]]), np.array([[lx / 3, lx / 2]])
    shot = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    seis = o.acoustic_vd_forward(np.full((n, n), c0), np.full((n, n), rho0), (dx, dx), dt, nt, halo, rcoef, False,
                                 [shot])[0][:, 0]
    times = dt / 2 + dt * np.arange(nt)
    dist = np.linalg.norm(src[0] - rec[0])
    G = np.zeros(nt)
    m = times - dist / c0 >= 0
    G[m] = 1.0 / (2 * np.pi * c0 ** 2 * np.sqrt(times[m] ** 2 - dist ** 2 / c0 ** 2))
    Gc = np.convolve(G, c0 ** 2 * rho0 * o.rickerstf(times, t0, f0) * dt)[:nt]
    err = np.trapezoid(np.abs(seis - Gc), times)
    assert err <= np.max(np.abs(Gc)) * 0.01 * (dt * nt)


def _problem(o, freetop):
    nx, nz, h, dt, nt, halo, f0 = 60, 50, 10.0, 1e-3, 300, 8, 10.0
    rng = np.random.default_rng(0)
    vp = 2000 + 200 * rng.random((nx, nz))
    rho = 2000 + 300 * rng.random((nx, nz))
    vp0 = np.full((nx, nz), 2000.0)
    vp0[2, 2] = 2300.0          # max(vp), which sets the C-PML profile, away from probed cells
    rho0 = np.full((nx, nz), 2100.0)
    tf = o.rickerstf(np.arange(nt) * dt, 1.2 / f0, f0)[:, None]
    src = np.array([[30 * h, 15 * h]])
    rec = np.stack([np.arange(12, 48, 3) * h, np.full(12, 35 * h)], 1)
    sh = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    sh["observed"] = o.acoustic_vd_forward(vp, rho, (h, h), dt, nt, halo, 1e-4, freetop, [sh])[0]
    return dict(vp0=vp0, rho0=rho0, h=h, dt=dt, nt=nt, halo=halo, sh=sh, freetop=freetop)


def test_vd_checkpointing_is_exact(oracle_mod):
    o = oracle_mod
    p = _problem(o, True)
    args = (p["vp0"], p["rho0"], (p["h"], p["h"]), p["dt"], p["nt"], p["halo"], 1e-4, True, [p["sh"]])
    g1, m1, _ = o.acoustic_vd_gradient(*args, check_freq=1)
    for cf in (3, 23, 100):
        g2, m2, _ = o.acoustic_vd_gradient(*args, check_freq=cf)
        assert np.array_equal(g1["vp"], g2["vp"]) and np.array_equal(g1["rho"], g2["rho"])
        assert m1 == m2


@pytest.mark.parametrize("freetop", [False, True])
def test_vd_gradient_matches_central_fd(oracle_mod, freetop):
    """The reference's "vp" and "rho" gradients equal -dJ/dvp and -dJ/drho
    of the discrete L2 misfit outside the C-PML (source cell included)."""
    o = oracle_mod
    p = _problem(o, freetop)
    sp = (p["h"], p["h"])
    g, _, _ = o.acoustic_vd_gradient(p["vp0"], p["rho0"], sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop, [p["sh"]],
                                     check_freq=20)

    def J(v, r):
        syn = o.acoustic_vd_forward(v, r, sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop, [p["sh"]])[0]
        return o.l2_misfit(syn, p["sh"]["observed"])

    for (i, j) in [(20, 25), (40, 30), (25, 40), (30, 15)]:
        d = 1e-2
        for key in ("vp", "rho"):
            a, b = p["vp0"].copy(), p["rho0"].copy()
            arr = a if key == "vp" else b
            arr[i, j] += d
            jp = J(a, b)
            arr[i, j] -= 2 * d
            jm = J(a, b)
            fd = (jp - jm) / (2 * d)
            assert abs(fd) > 0
            assert np.isclose(g[key][i, j], -fd, rtol=1e-5), (key, i, j, g[key][i, j], fd)


def test_vd_interpolation_matches_product_definition(oracle_mod):
    """Oracle staggered averaging (interpolations.jl:35-43) on the values of
    test_interpolations.jl."""
    o = oracle_mod
    m = np.array([1.0, 2.0, 3.0, 4.0])
    assert np.array_equal(o.interp_stag(m, 1), [1.5, 2.5, 3.5])
    np.testing.assert_allclose(o.interp_stag(m, 1, "harmonic"), [1.333, 2.4, 3.428], atol=1e-3)
