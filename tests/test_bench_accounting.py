"""bench.py's algorithmic-byte model (the numerator of every roofline figure)
against the per-cell counts of the step kernels: bytes per cell-step, e bytes
per value, halo 0 (the C-PML memory variables then reduce to the 2h+2 = 2
edge cells of each xi array per axis face, counted per launch as `strip`).

VD: forward 9e (r p, vx, vz, m0, m1x, m1z; w p, vx, vz), 10e with the
pressure history stored; adjoint 17e (9e + two history levels + three
gradient fields read and written).  CD: forward 4e (r p^n, p^n-1, fact;
w p^n+1), 5e with the replay history stored; adjoint 7e (+ r history, grad
read and written), 4e for the first step of each segment, whose correlation
runs in a kernel of its own.  Elastic, priced as the step with the stresses
kept on chip: forward 11e (r ux, uz at two levels, lambda, mu, mu_ihjh and
the two dt^2/rho fields; w ux, uz), 13e with the history stored; adjoint 27e
(11e + three history levels of two components + five gradient fields read
and written); the eight C-PML memory variables read and written per step."""
import bench

E = 4
N = 1000 * 100
NS = (1000, 100)
STRIP = bench._strip_cells(NS, 0) * 2 * E   # read and written once per launch


def test_vd_forward_and_adjoint_bytes():
    # nt a multiple of S: no final segment stored during the forward, every
    # segment replayed with S - 1 steps (each storing history)
    nt, S = 500, 50
    b = bench.alg_bytes("vd2d", NS, 0, E, [nt], S)
    rep = (nt // S) * (S - 1)
    assert b[0] == (nt * 9 * E * N + rep * 10 * E * N + (nt + rep) * STRIP, nt + rep)
    assert b[1] == (nt * 17 * E * N + nt * STRIP, nt)
    # a final partial segment is stored during the forward (10e per step)
    nt = 520
    b = bench.alg_bytes("vd2d", NS, 0, E, [nt], S)
    last, rep = 20, 10 * (S - 1)
    assert b[0] == ((nt - last) * 9 * E * N + last * 10 * E * N + rep * 10 * E * N + (nt + rep) * STRIP, nt + rep)


def test_vd_shot_group_bytes():
    """Shot groups of the block order (code + 1000 k, 2^k shots) read the
    model factors (3e) once per group and step: per shot 6e forward (+1e
    history) and 14e adjoint, plus 3e per group for each step of the
    group's longest shot."""
    nts, S = [500, 450, 520, 480], 50
    b1 = bench.alg_bytes("vd2d", NS, 0, E, nts, S, (316, 316))
    steps = [nt + (nt // S) * (S - 1) for nt in nts]
    for k, code in ((1, 1316), (2, 2316)):
        g = 1 << k
        b = bench.alg_bytes("vd2d", NS, 0, E, nts, S, (code, code))
        model_f = sum(3 * E * N * max(steps[i:i + g]) for i in range(0, len(nts), g))
        model_a = sum(3 * E * N * max(nts[i:i + g]) for i in range(0, len(nts), g))
        assert b[0][0] == b1[0][0] - sum(3 * E * N * st for st in steps) + model_f
        assert b[1][0] == b1[1][0] - sum(3 * E * N * nt for nt in nts) + model_a
        assert b[0][1] == b1[0][1] and b[1][1] == b1[1][1]
    # a group of two is priced as two shots per block (4xx)
    assert bench.alg_bytes("vd2d", NS, 0, E, nts, S, (1316, 1316)) == bench.alg_bytes("vd2d", NS, 0, E, nts, S,
                                                                                        (416, 416))


def test_cd2d_two_levels_per_launch_bytes():
    """Two CD levels per forward / replay launch read p^n, p^n-1 and fact
    once per launch (3e) and write each level and its D: per pair of levels
    5e (+2e D) instead of 8e (+2e); the adjoint is unchanged."""
    nt, S = 501, 50
    kl = (nt - 1) // S
    last, rep = nt - kl * S, kl * S
    one = bench.alg_bytes("cd2d", NS, 0, E, [nt], S, (2, 11))
    two = bench.alg_bytes("cd2d", NS, 0, E, [nt], S, (2, 12))
    nl = (nt + 1) // 2 + kl * (S // 2)
    assert two[0] == (nl * 3 * E * N + (nt + rep) * E * N + (last + rep) * E * N + (nt + rep) * STRIP, nl)
    assert two[1] == one[1]
    assert one[0][1] == nt + rep and two[0][0] < one[0][0]


def test_cd2d_two_adjoint_levels_per_launch_bytes():
    """Two CD adjoint levels per launch: per pair 3e read + 2e written, the
    gradient read and written once (2e) and one history level per
    correlating level; the top level of each segment does not correlate and
    an odd segment ends with a one-level launch (7e)."""
    S = 5
    for nt in (10, 11, 13):
        one = bench.alg_bytes("cd2d", NS, 0, E, [nt], S, (2, 11))
        two = bench.alg_bytes("cd2d", NS, 0, E, [nt], S, (2, 21))
        want_b = want_n = 0
        for k in range((nt - 1) // S + 1):
            L = min((k + 1) * S, nt) - k * S
            pairs, lone = L // 2, L % 2
            want_n += pairs + lone
            # the top pair correlates one level, the others two
            want_b += pairs * 7 * E * N + max(2 * pairs - 1, 0) * E * N + pairs * 2 * STRIP
            want_b += lone * ((7 if L > 1 else 4) * E * N + STRIP)
        assert two[1] == (want_b, want_n), nt
        assert two[0] == one[0]
        assert two[1][0] < one[1][0]


def test_cd_forward_and_adjoint_bytes():
    nt, S = 501, 50
    b = bench.alg_bytes("cd2d", NS, 0, E, [nt], S)
    kl = (nt - 1) // S
    last = nt - kl * S
    rep = nt - last
    assert b[0] == ((nt - last) * 4 * E * N + last * 5 * E * N + rep * 5 * E * N + (nt + rep) * STRIP, nt + rep)
    nseg = kl + 1
    assert b[1] == ((nt - nseg) * 7 * E * N + nseg * 4 * E * N + nt * STRIP, nt)


def test_elastic_bytes_per_step():
    nt, S = 520, 50
    b = bench.alg_bytes("el2d", NS, 0, E, [nt], S, variant=(1, 1))
    last, rep = 20, (nt // S) * (S - 1)
    est = (4 * NS[1] + 4 * NS[0]) * 2 * E          # halo 0: 4 rows of 2h+2 per axis
    # two launches per step, stored and replayed steps at 13e
    assert b[0] == ((nt - last) * 11 * E * N + (last + rep) * 13 * E * N + (nt + rep) * est, 2 * (nt + rep))
    assert b[1] == (nt * 27 * E * N + nt * est, 2 * nt)
    # one launch per step (z-marching variant): same bytes, half the launches
    m = bench.alg_bytes("el2d", NS, 0, E, [nt], S, variant=(104, 104))
    assert m[0] == (b[0][0], nt + rep) and m[1] == (b[1][0], nt)


def test_shots_add_up():
    one = bench.alg_bytes("vd2d", NS, 20, E, [300], 50)
    two = bench.alg_bytes("vd2d", NS, 20, E, [300, 300], 50)
    assert two[0] == (2 * one[0][0], 2 * one[0][1])
    assert two[1] == (2 * one[1][0], 2 * one[1][1])


def test_strips_counted_once_per_launch():
    # the C-PML strips add the same bytes to every launch of a family
    nt, S = 100, 50
    b0 = bench.alg_bytes("cd2d", NS, 0, E, [nt], S)
    b1 = bench.alg_bytes("cd2d", NS, 10, E, [nt], S)
    extra = (bench._strip_cells(NS, 10) - bench._strip_cells(NS, 0)) * 2 * E
    assert b1[0][0] - b0[0][0] == b0[0][1] * extra
    assert b1[1][0] - b0[1][0] == b0[1][1] * extra


def test_vd_pair_variant_reads_model_once_per_pair():
    # two shots per block (variant 416): m0, m1x, m1z (3e) are read once per
    # shot pair and step (the pair steps as long as its longer shot)
    nts, S = [500, 420, 333], 50
    one = bench.alg_bytes("vd2d", NS, 0, E, nts, S, variant=(316, 316))
    rep = [(nt // S) * (S - 1) for nt in nts]
    for v, fam, saved in (((416, 316), 0, 3 * E * N * (sum(nt + r for nt, r in zip(nts, rep))
                                                         - max(500 + rep[0], 420 + rep[1]) - (333 + rep[2]))),
                          ((316, 416), 1, 3 * E * N * (sum(nts) - 500 - 333))):
        pair = bench.alg_bytes("vd2d", NS, 0, E, nts, S, variant=v)
        assert one[fam][0] - pair[fam][0] == saved
        assert one[1 - fam][0] == pair[1 - fam][0]
