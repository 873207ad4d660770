"""Pins the 1D part of the CPU oracle (acoustic1D_xPU.jl,
acoustic1D_VD_xPU.jl): the reference's own 1D analytic tests with their 1%
integrated-error bound, checkpointed == full storage bitwise, central
finite differences of the misfit for the "vp" (CD) and "vp"/"rho" (VD)
gradients, and the edge closed form of the CD rotation rule."""
import numpy as np
import pytest


def _greens_1d(c0, times, dist):
    """1D Green's function H(t - |x|/c) / (2c) (setup_models.jl:181-199)."""
    G = np.zeros(len(times))
    G[times - dist / c0 >= 0] = 1.0 / (2 * c0)
    return G


@pytest.mark.parametrize("halo,rcoef,nt", [(0, 1.0, 500), (20, 1e-4, 1000)])
def test_cd1d_forward_analytic(oracle_mod, halo, rcoef, nt):
    """test_analytical_vs_numerical_acoustic_constant_density.jl:18-62 with
    setup_constant_vel_1D_CPML and analytical_solution_constant_vel_1D
    (setup_models.jl:24-50, :181-199)."""
    o = oracle_mod
    c0, nx, dx, f0 = 1000.0, 501, 2.5, 5.0
    dt = 0.99 * dx / c0
    lx = (nx - 1) * dx
    tf = o.rickerstf(dt * np.arange(nt), 2 / f0, f0)[:, None]
    src, rec = np.array([[lx / 2]]), np.array([[lx / 3]])
    shot = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    seis = o.acoustic_cd1d_forward(np.full(nx, c0), (dx,), dt, nt, halo, rcoef, False, [shot])[0][:, 0]
    times = dt * np.arange(1, nt + 1)
    Gc = np.convolve(_greens_1d(c0, times, abs(lx / 2 - lx / 3)), c0 ** 2 * tf[:, 0] * dt)[:nt]
    err = np.trapezoid(np.abs(seis - Gc), times)
    assert np.max(np.abs(Gc)) > 0
    assert err <= np.max(np.abs(Gc)) * 0.01 * (dt * nt)


@pytest.mark.parametrize("halo,rcoef,nt", [(0, 1.0, 500), (20, 1e-4, 5000)])
def test_vd1d_forward_analytic(oracle_mod, halo, rcoef, nt):
    """test_analytical_vs_numerical_acoustic_variable_density.jl:18-65 with
    setup_constant_vel_rho_1D_CPML and analytical_solution_constant_vel_
    constant_density_1D (setup_models.jl:52-74, :200-218): Gaussian-
    derivative source, half-step time axis."""
    o = oracle_mod
    c0, rho0, nx, dx, f0 = 1000.0, 1500.0, 501, 2.5, 5.0
    dt = 0.99 * dx / c0 * 6 / 7
    t0 = 2 / f0
    lx = (nx - 1) * dx
    tf = o.gaussderivstf(dt * np.arange(nt), t0, f0)[:, None]
    src, rec = np.array([[lx / 2]]), np.array([[lx / 3]])
    shot = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    seis = o.acoustic_vd1d_forward(np.full(nx, c0), np.full(nx, rho0), (dx,), dt, nt, halo, rcoef, False,
                                   [shot])[0][:, 0]
    times = dt / 2 + dt * np.arange(nt)
    Gc = np.convolve(_greens_1d(c0, times, abs(lx / 2 - lx / 3)),
                     c0 ** 2 * rho0 * o.rickerstf(times, t0, f0) * dt)[:nt]
    err = np.trapezoid(np.abs(seis - Gc), times)
    assert np.max(np.abs(Gc)) > 0
    assert err <= np.max(np.abs(Gc)) * 0.01 * (dt * nt)


def _problem(o, vd, freetop, edge=False):
    nx, h, dt, nt, halo, f0 = 120, 10.0, 1e-3, 400, 10, 10.0
    rng = np.random.default_rng(5)
    vp = 2000 + 200 * rng.random(nx)
    rho = 2000 + 300 * rng.random(nx)
    vp0 = np.full(nx, 2000.0)
    vp0[2] = 2300.0          # max(vp), which sets the C-PML profile, away from probed cells
    rho0 = np.full(nx, 2100.0)
    tf = o.rickerstf(np.arange(nt) * dt, 1.2 / f0, f0)[:, None]
    src = np.array([[0.0 if edge else 40 * h]])
    rec = np.array([[x * h] for x in (25, 55, 70, 90)] + ([[0.0], [(nx - 1) * h]] if edge else []))
    sh = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    if vd:
        sh["observed"] = o.acoustic_vd1d_forward(vp, rho, (h,), dt, nt, halo, 1e-4, freetop, [sh])[0]
    else:
        sh["observed"] = o.acoustic_cd1d_forward(vp, (h,), dt, nt, halo, 1e-4, freetop, [sh])[0]
    return dict(vp0=vp0, rho0=rho0, h=h, dt=dt, nt=nt, halo=halo, sh=sh, freetop=freetop)


@pytest.mark.parametrize("vd", [False, True])
@pytest.mark.parametrize("edge", [False, True])
def test_1d_checkpointing_is_exact(oracle_mod, vd, edge):
    o = oracle_mod
    p = _problem(o, vd, True, edge)
    if vd:
        run = lambda cf: o.acoustic_vd1d_gradient(p["vp0"], p["rho0"], (p["h"],), p["dt"], p["nt"], p["halo"], 1e-4,
                                                  True, [p["sh"]], check_freq=cf)
    else:
        run = lambda cf: o.acoustic_cd1d_gradient(p["vp0"], (p["h"],), p["dt"], p["nt"], p["halo"], 1e-4, True,
                                                  [p["sh"]], check_freq=cf)
    g1, m1, _ = run(1)
    assert np.max(np.abs(g1["vp"])) > 0
    for cf in (3, 23, 100, 399):
        g2, m2, _ = run(cf)
        for k in g1:
            assert np.array_equal(g1[k], g2[k]), (k, cf)
        assert m1 == m2


@pytest.mark.parametrize("freetop", [False, True])
def test_cd1d_gradient_matches_central_fd(oracle_mod, freetop):
    """The "vp" gradient equals -dJ/dvp of the discrete L2 misfit outside
    the C-PML (source cell included)."""
    o = oracle_mod
    p = _problem(o, False, freetop)
    sp = (p["h"],)
    g, _, _ = o.acoustic_cd1d_gradient(p["vp0"], sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop, [p["sh"]],
                                       check_freq=20)

    def J(v):
        syn = o.acoustic_cd1d_forward(v, sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop, [p["sh"]])[0]
        return o.l2_misfit(syn, p["sh"]["observed"])

    for i in (15, 33, 40, 60, 85):
        d = 1e-2
        a = p["vp0"].copy()
        a[i] += d
        jp = J(a)
        a[i] -= 2 * d
        jm = J(a)
        fd = (jp - jm) / (2 * d)
        assert abs(fd) > 0
        assert np.isclose(g["vp"][i], -fd, rtol=1e-6), (i, g["vp"][i], fd)


@pytest.mark.parametrize("freetop", [False, True])
def test_vd1d_gradient_matches_central_fd(oracle_mod, freetop):
    """"vp" and "rho" gradients equal -dJ/dvp and -dJ/drho, inside 3:nx-3
    (correlate_gradient_m1! in 1D skips the first and last staggered cells,
    acoustic1D_VD_xPU.jl:130-134, so rho there is not the full derivative)."""
    o = oracle_mod
    p = _problem(o, True, freetop)
    sp = (p["h"],)
    g, _, _ = o.acoustic_vd1d_gradient(p["vp0"], p["rho0"], sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop,
                                       [p["sh"]], check_freq=20)

    def J(v, r):
        syn = o.acoustic_vd1d_forward(v, r, sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop, [p["sh"]])[0]
        return o.l2_misfit(syn, p["sh"]["observed"])

    for i in (15, 33, 40, 60, 85):
        d = 1e-2
        for key in ("vp", "rho"):
            a, b = p["vp0"].copy(), p["rho0"].copy()
            arr = a if key == "vp" else b
            arr[i] += d
            jp = J(a, b)
            arr[i] -= 2 * d
            jm = J(a, b)
            fd = (jp - jm) / (2 * d)
            assert abs(fd) > 0
            assert np.isclose(g[key][i], -fd, rtol=1e-5), (key, i, g[key][i], fd)


def test_cd1d_edge_cell_keeps_level_it_minus_2(oracle_mod):
    """acoustic1D_xPU.jl:89-92 binds pnew to pold's array: the trace at an
    edge source cell is sum_{j = k mod 2, j <= k} s(j)."""
    o = oracle_mod
    nx, nt = 60, 80
    rng = np.random.default_rng(9)
    fact = rng.uniform(3.0, 5.0, nx) * 1e-2
    cpml = o.init_bdc(np.sqrt(fact) / 1e-3, 1e-3, (10.0,), 5, 1e-4, False, 15.0)
    stf = rng.standard_normal((nt, 2))
    src = np.array([[1], [nx]])
    rec = np.array([[1], [nx], [30]])
    tr = o.cd1d_forward(fact, (10.0,), 1e-3, 5, cpml, nt, src, stf, rec)
    for k in range(2):
        want = np.zeros(nt)
        for t in range(nt):
            want[t] = (want[t - 2] if t >= 2 else 0.0) + stf[t, k]
        np.testing.assert_allclose(tr[:, k], want, rtol=1e-13, atol=0)
    assert np.abs(tr[:, 2]).max() > 0


def test_vd1d_m1_correlation_range(oracle_mod):
    """correlate_gradient_m1! in 1D runs over 2:nx-2 (acoustic1D_VD_xPU.jl:
    130-134): the raw staggered gradient is zero on the first and last of the
    nx-1 staggered cells and not on their neighbours."""
    o = oracle_mod
    nx, nt, h, dt = 40, 200, 10.0, 1e-3
    rng = np.random.default_rng(4)
    m0 = rng.uniform(0.8, 1.2, nx) * 8e3
    m1 = rng.uniform(0.8, 1.2, nx - 1) * 5e-7
    cpml = o.init_bdc(np.full(nx, 2000.0), dt, (h,), 0, 1.0, False, 15.0)
    stf = rng.standard_normal((nt, 1))
    adj = rng.standard_normal((nt, 2))
    (g0, g1), _ = o.vd1d_gradient(m0, m1, (h,), dt, 0, cpml, nt, np.array([[2]]), stf, np.array([[3], [nx - 1]]),
                                  adj, 7)
    assert g1[0] == 0 and g1[nx - 2] == 0
    assert g1[1] != 0 and g1[nx - 3] != 0
