"""The oracle's CD time step on the outermost grid cells.

The reference rotates its three pressure fields with three sequential
assignments (acoustic2D_xPU.jl:122-124, :165-167 for the adjoint;
acoustic3D_xPU.jl:155-157, :214-216):

    pold <- pcur;  pcur <- pnew;  pnew <- pold

The third assignment reads the *new* pold, so after the first step pnew and
pold are the same array and the time loop runs on two arrays.  The stencil
(update_p_CPML!) covers 2:n-1 only, so a cell on the outer rows, columns or
faces keeps the value of level it-2 plus whatever is injected there:

    p^k[c] = p^{k-2}[c] + s(k),        p^0 = p^{-1} = 0       (forward)
    a^k[c] = a^{k+2}[c] + fact[c] r(k), a^{nt+1} = a^{nt+2} = 0 (adjoint)

independent of everything else in the grid.  These closed forms pin the
oracle's edge rule: a three-array rotation (level it-3 on the edges) or a
zeroed edge fails them.  Because of the rule, a checkpoint replay
reproduces the forward exactly, so the gradient cannot depend on the
checkpoint interval; that is checked bitwise with sources and receivers on
the edges.
"""
import numpy as np
import pytest

H, DT = 10.0, 1e-3


def _edge_series(s):
    """p^k = p^{k-2} + s(k), k = 1..nt (returned 0-based: out[k-1] = p^k)."""
    out = np.zeros_like(s)
    for k in range(len(s)):
        out[k] = (out[k - 2] if k >= 2 else 0.0) + s[k]
    return out


def _adj_series(r):
    """a^k = a^{k+2} + r(k) for k = nt..1 (out[k-1] = a^k)."""
    nt = len(r)
    out = np.zeros_like(r)
    for k in range(nt - 1, -1, -1):
        out[k] = (out[k + 2] if k + 2 < nt else 0.0) + r[k]
    return out


def _grad_edge(p, a):
    """sum_it a^it (p^{it-2} - 2 p^{it-1} + p^it) / dt^2 at one cell."""
    pe = np.concatenate([[0.0, 0.0], p])   # pe[k+1] = p^k, k = -1..nt
    g = 0.0
    for it in range(1, len(p) + 1):
        g += a[it - 1] * (pe[it - 1] - 2 * pe[it] + pe[it + 1])
    return g / DT ** 2


def _case(o, dims, seed):
    rng = np.random.default_rng(seed)
    n = {2: (40, 34), 3: (22, 20, 18)}[dims]
    vp = rng.uniform(1800.0, 2600.0, n)
    fact = vp ** 2 * DT ** 2
    cpml = o.init_bdc(vp, DT, (H,) * dims, 5, 1e-4, dims == 2, 15.0)
    if dims == 2:
        edge = np.array([[1, 1], [17, 1], [n[0], 9], [6, n[1]], [1, 20]])
        inner = np.array([[20, 17]])
    else:
        edge = np.array([[1, 1, 1], [9, 1, 7], [n[0], 8, 9], [5, 11, n[2]], [1, 6, 4]])
        inner = np.array([[11, 10, 9]])
    nt = 90
    stf = rng.standard_normal((nt, len(edge))) * 1e-3
    return dict(n=n, fact=fact, cpml=cpml, src=edge, rec=np.concatenate([edge, inner]), stf=stf, nt=nt)


@pytest.mark.parametrize("dims", [2, 3])
def test_edge_cells_keep_level_it_minus_2(oracle_mod, dims):
    c = _case(oracle_mod, dims, 11)
    fwd = oracle_mod.cd2d_forward if dims == 2 else oracle_mod.cd3d_forward
    tr = fwd(c["fact"], (H,) * dims, DT, 5, c["cpml"], c["nt"], c["src"], c["stf"], c["rec"])
    for k in range(len(c["src"])):
        np.testing.assert_allclose(tr[:, k], _edge_series(c["stf"][:, k]), rtol=1e-13, atol=0)
    # the interior receiver sees a wave, which a three-array rotation would
    # change as well (its neighbours read the edge values)
    assert np.abs(tr[:, -1]).max() > 0


@pytest.mark.parametrize("dims", [2, 3])
def test_edge_gradient_closed_form(oracle_mod, dims):
    c = _case(oracle_mod, dims, 12)
    rng = np.random.default_rng(13)
    adj = rng.standard_normal((c["nt"], len(c["rec"])))
    grad_fn = oracle_mod.cd2d_gradient if dims == 2 else oracle_mod.cd3d_gradient
    g, tr = grad_fn(c["fact"], (H,) * dims, DT, 5, c["cpml"], c["nt"], c["src"], c["stf"], c["rec"], adj, 7)
    for k, cell in enumerate(c["src"]):
        p = _edge_series(c["stf"][:, k])
        a = _adj_series(adj[:, k] * c["fact"][tuple(cell - 1)])
        np.testing.assert_allclose(g[tuple(cell - 1)], _grad_edge(p, a), rtol=1e-11)


@pytest.mark.parametrize("dims", [2, 3])
def test_edge_points_checkpointing_is_exact(oracle_mod, dims):
    c = _case(oracle_mod, dims, 14)
    rng = np.random.default_rng(15)
    adj = rng.standard_normal((c["nt"], len(c["rec"])))
    grad_fn = oracle_mod.cd2d_gradient if dims == 2 else oracle_mod.cd3d_gradient
    gs = [grad_fn(c["fact"], (H,) * dims, DT, 5, c["cpml"], c["nt"], c["src"], c["stf"], c["rec"], adj, cf)[0]
          for cf in (1, 7, 30, 89)]
    assert np.abs(gs[0]).max() > 0
    for g in gs[1:]:
        assert np.array_equal(g, gs[0])
