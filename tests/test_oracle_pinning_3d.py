"""Pins the 3D constant-density part of the CPU oracle: the reference's own
3D analytic test and tolerance, checkpoint exactness, and central finite
differences of the misfit for the vp gradient."""
import numpy as np
import pytest


@pytest.mark.parametrize("halo,rcoef,nt", [(0, 1.0, 200), (20, 1e-4, 400)])
def test_cd3d_forward_analytic(oracle_mod, halo, rcoef, nt):
    """test_analytical_vs_numerical_acoustic_constant_density.jl:108-150 with
    setup_models.jl:155-179 and analytical_solution_constant_vel_3D
    (setup_models.jl:260-280): delta Green's function, 1% bound."""
    o = oracle_mod
    c0, n, dx, f0 = 1000.0, 121, 8.0, 5.0
    dt = 0.99 * dx / c0 / np.sqrt(3)
    L = (n - 1) * dx
    tf = o.rickerstf(dt * np.arange(nt), 2 / f0, f0)[:, None]
    src, rec = np.array([[L / 2, L / 2, L / 2]]), np.array([[L / 3, L / 2, L / 2]])
    sh = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    seis = o.acoustic_cd3d_forward(np.full((n, n, n), c0), (dx, dx, dx), dt, nt, halo, rcoef, False, [sh])[0][:, 0]
    times = dt * np.arange(1, nt + 1)
    dist = np.linalg.norm(src[0] - rec[0])
    G = np.zeros(nt)
    first = np.nonzero(times - dist / c0 >= 0)[0][0]
    G[first] = 1.0 / (4 * np.pi * c0 ** 2 * dist)
    Gc = np.convolve(G, c0 ** 2 * tf[:, 0])[:nt]
    err = np.trapezoid(np.abs(seis - Gc), times)
    assert err <= np.max(np.abs(Gc)) * 0.01 * (dt * nt)


def _problem(o, freetop):
    n, h, dt, nt, halo, f0 = (28, 24, 22), 10.0, 1e-3, 160, 5, 10.0
    rng = np.random.default_rng(4)
    vp = 2000 + 200 * rng.random(n)
    vp0 = np.full(n, 2000.0)
    vp0[1, 1, 1] = 2300.0      # max(vp), which sets the C-PML profile, away from probed cells
    tf = o.rickerstf(np.arange(nt) * dt, 1.2 / f0, f0)[:, None]
    src = np.array([[14 * h, 12 * h, 6 * h]])
    rec = np.array([[x * h, y * h, 15 * h] for x in (8, 12, 16, 20) for y in (8, 14)])
    sh = dict(src_pos=src, srctf=tf, f0=f0, rec_pos=rec)
    sh["observed"] = o.acoustic_cd3d_forward(vp, (h, h, h), dt, nt, halo, 1e-4, freetop, [sh])[0]
    return dict(vp0=vp0, h=h, dt=dt, nt=nt, halo=halo, sh=sh)


def test_cd3d_checkpointing_is_exact(oracle_mod):
    o = oracle_mod
    p = _problem(o, True)
    args = (p["vp0"], (p["h"],) * 3, p["dt"], p["nt"], p["halo"], 1e-4, True, [p["sh"]])
    g1, m1, _ = o.acoustic_cd3d_gradient(*args, check_freq=1)
    for cf in (7, 40):
        g2, m2, _ = o.acoustic_cd3d_gradient(*args, check_freq=cf)
        assert np.array_equal(g1["vp"], g2["vp"]) and m1 == m2


@pytest.mark.parametrize("freetop", [False, True])
def test_cd3d_gradient_matches_central_fd(oracle_mod, freetop):
    o = oracle_mod
    p = _problem(o, freetop)
    sp = (p["h"],) * 3
    g, _, _ = o.acoustic_cd3d_gradient(p["vp0"], sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop, [p["sh"]],
                                       check_freq=20)

    def J(v):
        syn = o.acoustic_cd3d_forward(v, sp, p["dt"], p["nt"], p["halo"], 1e-4, freetop, [p["sh"]])[0]
        return o.l2_misfit(syn, p["sh"]["observed"])

    for idx in [(10, 10, 10), (17, 13, 12), (14, 12, 6), (12, 9, 14)]:
        d = 1e-2
        a, b = p["vp0"].copy(), p["vp0"].copy()
        a[idx] += d
        b[idx] -= d
        fd = (J(a) - J(b)) / (2 * d)
        assert abs(fd) > 0
        assert np.isclose(g["vp"][idx], -fd, rtol=1e-5), (idx, g["vp"][idx], fd)
