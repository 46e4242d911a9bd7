"""Pins the CPU oracle against closed forms, invariants, special cases and
brute force (the reference ships no golden vectors; its only known-answer
check is AbstractEvaluator::assessMeasureValidy, Evaluators/AbstractEvaluator.cpp:281-338).
Each check is chosen so that a plausible slip in the restatement (a dropped
term, a transposed operand, a wrong index or sign) fails it."""
import os
import numpy as np
import pytest

from gpumesh_amd import cube_mesh

S3 = np.sqrt(3.0)
REG_TET = np.array([[0, 0, 0], [1, 0, 0], [0.5, S3 / 2, 0], [0.5, S3 / 6, np.sqrt(2.0 / 3)]])
UNIT_HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
REG_PRISM_FREE = None


@pytest.fixture(scope="module")
def om(oracle):
    m = cube_mesh(2, "tet", 0.0, seed=1)
    o = oracle.OracleMesh.from_cube(m)
    o.set_crew("mean_ratio", "uniform")
    return o


def mean_ratio_closed_form(p):
    """Mean ratio of a tetrahedron from volume and squared edge lengths:
    q = 12 (3|V|)^(2/3) / sum l_ij^2, signed by the orientation
    det(v3-v0, v3-v1, v3-v2)."""
    p = np.asarray(p, float)
    d = np.linalg.det(np.stack([p[3] - p[0], p[3] - p[1], p[3] - p[2]], axis=1))
    vol = abs(d) / 6.0
    l2 = sum(np.sum((p[i] - p[j]) ** 2) for i in range(4) for j in range(i + 1, 4))
    return np.sign(d) * 12.0 * (3.0 * vol) ** (2.0 / 3.0) / l2


def test_reference_known_answers(om):
    # assessMeasureValidy: regular tet and hex score 1 within 1e-6.
    assert abs(om.tet_quality(REG_TET) - 1.0) < 1e-12
    assert abs(om.hex_quality(UNIT_HEX) - 1.0) < 1e-12


def test_mean_ratio_special_tets(om):
    assert om.tet_quality(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])) == pytest.approx(0.839947366596582, abs=1e-15)
    kuhn = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1.0]])
    assert om.tet_quality(kuhn) == pytest.approx(0.7559526299369238, abs=1e-15)
    inverted = REG_TET[[1, 0, 2, 3]]
    assert om.tet_quality(inverted) == pytest.approx(-1.0, abs=1e-12)
    flat = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.0]])
    assert om.tet_quality(flat) == 0.0


def test_mean_ratio_matches_closed_form_on_random_tets(om):
    rng = np.random.default_rng(3)
    for _ in range(300):
        p = rng.normal(size=(4, 3))
        assert om.tet_quality(p) == pytest.approx(mean_ratio_closed_form(p), rel=1e-12, abs=1e-13)


def test_mean_ratio_invariances(om):
    rng = np.random.default_rng(4)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    for _ in range(50):
        p = rng.normal(size=(4, 3))
        base = om.tet_quality(p)
        assert om.tet_quality(p * 3.7 + 1.5) == pytest.approx(base, rel=1e-12)
        assert om.tet_quality(p @ q.T) == pytest.approx(base, rel=1e-12)
        h = UNIT_HEX + rng.uniform(-0.2, 0.2, size=UNIT_HEX.shape)
        hb = om.hex_quality(h)
        assert om.hex_quality(h * 0.25 - 2.0) == pytest.approx(hb, rel=1e-12)
        assert om.hex_quality(h @ q.T) == pytest.approx(hb, rel=1e-12)


def test_hex_boxes_and_affine_images(om):
    # A box a x b x c: every corner has q = 3 (abc)^(2/3) / (a^2 + b^2 + c^2).
    for a, b, c in [(1, 1, 2), (0.5, 2, 3), (1, 1, 1)]:
        box = UNIT_HEX * np.array([a, b, c])
        expect = 3 * (a * b * c) ** (2 / 3) / (a * a + b * b + c * c)
        assert om.hex_quality(box) == pytest.approx(expect, rel=1e-13)
    assert om.hex_quality(UNIT_HEX * np.array([1, 1, 2])) == pytest.approx(2 ** (-1 / 3), rel=1e-15)
    # A parallelepiped (affine image of the cube): all 8 corner matrices equal J.
    J = np.array([[1.0, 0.3, 0.1], [0.0, 1.2, 0.2], [0.1, 0.0, 0.9]])
    ph = UNIT_HEX @ J.T
    expect = 3 * abs(np.linalg.det(J)) ** (2 / 3) / np.sum(J * J)
    assert om.hex_quality(ph) == pytest.approx(expect, rel=1e-12)
    # One inverted corner makes the hex quality the minimum corner value (<= 0).
    bad = UNIT_HEX.copy()
    bad[6] = [0.2, 0.2, 0.2]
    assert om.hex_quality(bad) <= 0.0


def test_volumes(oracle, om):
    assert om.tet_volume(REG_TET) == pytest.approx(np.sqrt(2) / 12, rel=1e-14)
    J = np.array([[1.0, 0.3, 0.1], [0.0, 1.2, 0.2], [0.1, 0.0, 0.9]])
    assert om.hex_volume(UNIT_HEX @ J.T) == pytest.approx(np.linalg.det(J), rel=1e-13)
    assert om.hex_volume(UNIT_HEX * 2.0) == pytest.approx(8.0, rel=1e-15)


def test_metric_conformity_known_values(oracle):
    m = cube_mesh(2, "tet", 0.0, seed=1)
    o = oracle.OracleMesh.from_cube(m)
    # Uniform metric K^2 I with K = 1 and a regular unit tet: Mk = I = Ms, T = 0, q = 1.
    o.set_crew("metric_conformity", "uniform", scaling=1.0)
    assert o.tet_quality(REG_TET) == pytest.approx(1.0, abs=1e-13)
    assert o.hex_quality(UNIT_HEX) == pytest.approx(1.0, abs=1e-13)
    # K = 2: Ms = 4I, Mk = I -> T = (4 + 1/4 - 2) I, |T|_F = sqrt(3) * 2.25.
    o.set_crew("metric_conformity", "uniform", scaling=2.0)
    assert o.tet_quality(REG_TET) == pytest.approx(1 / np.sqrt(1 + np.sqrt(3) * 2.25), rel=1e-13)
    # Inverted element: negative sign.
    assert o.tet_quality(REG_TET[[1, 0, 2, 3]]) < 0


def test_metric_free_and_uniform_metric_wise_distance(oracle):
    m = cube_mesh(2, "tet", 0.0, seed=1)
    o = oracle.OracleMesh.from_cube(m)
    a, b = np.array([0.1, 0.2, 0.3]), np.array([0.7, -0.1, 0.5])
    o.set_crew("mean_ratio", "uniform", scaling=3.0)
    assert o.riemannian_distance(a, b) == pytest.approx(3.0 * np.linalg.norm(b - a), rel=1e-15)
    assert np.allclose(o.metric_at(a), 9.0 * np.eye(3), rtol=0, atol=0)
    # Metric-wise subdivision of a constant metric integrates exactly.
    o.set_crew("mean_ratio", "uniform", measurer="metric_wise", scaling=3.0)
    assert o.riemannian_distance(a, b) == pytest.approx(3.0 * np.linalg.norm(b - a), rel=1e-14)


def test_shell_metric_eigenstructure(oracle):
    m = cube_mesh(2, "tet", 0.0, seed=1)
    o = oracle.OracleMesh.from_cube(m)
    K = 64.0
    o.set_crew("metric_conformity", "analytic", scaling=K)
    rng = np.random.default_rng(5)
    for _ in range(20):
        p = rng.uniform(0, 1, 3)
        d = p - 0.5
        r = np.linalg.norm(d)
        a = 1 + 9 * np.exp(-50 * (r - 0.3) ** 2)
        M = o.metric_at(p)
        assert np.allclose(M, M.T, rtol=1e-15, atol=0)
        rh = d / r
        assert np.allclose(M @ rh, K * K * rh, rtol=1e-11)
        t = np.cross(rh, [0.3, 0.5, 0.7])
        t /= np.linalg.norm(t)
        assert np.allclose(M @ t, K * K * a * a * t, rtol=1e-11)
    assert np.allclose(o.metric_at([0.5, 0.5, 0.5]), K * K * (1 + 9 * np.exp(-50 * 0.09)) ** 2 * np.eye(3), rtol=1e-13)


def test_texture_sampler_reproduces_affine_fields(oracle):
    """Trilinear interpolation is exact for affine fields between cell centres."""
    m = cube_mesh(2, "tet", 0.0, seed=1)
    o = oracle.OracleMesh.from_cube(m)
    n = 8
    size = np.array([n, n, n], np.int32)
    tmin = np.zeros(3)
    text = np.ones(3)
    cs = 1.0 / n
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)  # i,j,k
    centres = (idx + 0.5) * cs
    coef = np.array([[2.0, 1.0, 0.5, 4.0], [0.0, 1.0, 2.0, 1.0], [0.25, 0.0, 1.0, 0.5],
                     [1.0, 2.0, 0.0, 6.0], [0.5, 0.5, 0.5, 0.0], [0.0, 0.0, 3.0, 5.0]])

    def field(p):
        return np.array([c[0] * p[..., 0] + c[1] * p[..., 1] + c[2] * p[..., 2] + c[3] for c in coef])

    vals = field(centres)  # (6, i, j, k)
    texels = np.transpose(vals, (3, 2, 1, 0)).astype(np.float32)  # (k, j, i, 6)
    assert np.array_equal(texels.astype(np.float64), np.transpose(vals, (3, 2, 1, 0)))
    o.set_crew("metric_conformity", "texture", texture=(size, tmin, text, texels.reshape(-1)))
    rng = np.random.default_rng(6)
    for _ in range(50):
        p = rng.uniform(cs / 2, 1 - cs / 2, 3)
        M = o.metric_at(p)
        f = field(p)
        expect = np.array([[f[0], f[1], f[2]], [f[1], f[3], f[4]], [f[2], f[4], f[5]]])
        assert np.allclose(M, expect, rtol=1e-13, atol=1e-13)
    # Outside the centre range the sampler clamps to the boundary cell values.
    M = o.metric_at(np.array([0.0, 0.0, 0.0]))
    f = field(np.array([cs / 2, cs / 2, cs / 2]))
    assert M[0, 0] == pytest.approx(f[0], rel=1e-13)


def test_project_libm_accuracy(oracle):
    """gm_pow23 / gm_cbrt / gm_exp against extended-precision values.

    glibc's pow(x, 2.0/3.0) uses the exponent rounded to double, which is
    off from 2/3 by 3.7e-17 and so from x^(2/3) by |ln x| * 3.7e-17 relative;
    the project routine computes x^(2/3) itself.  Both are compared with the
    true value, and the two with each other within that known gap."""
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.uniform(0, 10, 20000), 10.0 ** rng.uniform(-300, 300, 20000),
                        np.array([5e-324, 1e-310, 2.2250738585072014e-308, 1.0, 8.0, 27.0, 1e308])])
    xl = x.astype(np.longdouble)
    true23 = (np.cbrt(xl) ** 2).astype(np.float64)
    true13 = np.cbrt(xl).astype(np.float64)
    p = oracle.gm_libm(2, x)
    assert np.max(np.abs(p - true23) / np.spacing(true23)) <= 1.0
    c = oracle.gm_libm(3, x)
    assert np.max(np.abs(c - true13) / np.spacing(true13)) <= 1.0
    glibc = x ** (2.0 / 3.0)
    gap = np.abs(np.log(x)) * 3.71e-17 * true23 / np.spacing(true23) + 3.0
    assert np.all(np.abs(p - glibc) / np.spacing(true23) <= gap)
    assert oracle.gm_libm(2, np.array([0.0]))[0] == 0.0
    assert np.isnan(oracle.gm_libm(2, np.array([-1.0]))[0])
    e = rng.uniform(-700, 700, 20000)
    pe = oracle.gm_libm(4, e)
    refe = np.exp(e.astype(np.longdouble)).astype(np.float64)
    assert np.max(np.abs(pe - refe) / np.spacing(refe)) <= 1.0
    assert oracle.gm_libm(4, np.array([0.0]))[0] == 1.0


def test_histogram_buckets(oracle):
    o = oracle.OracleMesh.from_cube(cube_mesh(1, "tet", 0.0, seed=1))
    r = o.evaluate()
    # The 6 tets of one Kuhn cube have q = 0.75595..., bucket 15.
    assert r["n"] == 6 and r["hist"][15] == 6 and r["hist"].sum() == 6
    assert r["min"] == pytest.approx(0.7559526299369238, abs=1e-15)
    assert r["mean"] == pytest.approx(0.7559526299369238, abs=1e-15)


def test_unjittered_grid_is_uniform(oracle):
    for kind, q in [("tet", 0.7559526299369238), ("hex", 1.0)]:
        o = oracle.OracleMesh.from_cube(cube_mesh(5, kind, 0.0, seed=1))
        r = o.evaluate()
        assert r["min"] == pytest.approx(q, abs=1e-14)
        assert r["mean"] == pytest.approx(q, abs=1e-13)


def test_colouring_invariants(oracle):
    for kind, n in [("tet", 7), ("hex", 9)]:
        m = cube_mesh(n, kind, 0.3, seed=2)
        o = oracle.OracleMesh.from_cube(m)
        groups, tr = o.groups()
        t, h = o.elements()
        elems = t if kind == "tet" else h
        V = m.nverts
        nfixed = int(m.fixed.sum())
        assert tr[0, 0] == 0 and tr[0, 1] == nfixed and tr[3, 1] == V   # interior end == V
        colour = -np.ones(V, int)
        covered = np.zeros(V, bool)
        last_end = nfixed
        for g, r in enumerate(groups):
            b, e = r[4], r[5]
            if e > b:
                assert b == last_end          # contiguous, in colour order
                last_end = e
                colour[b:e] = g
                covered[b:e] = True
        assert covered[nfixed:].all() and not covered[:nfixed].any()
        # No two same-colour movable vertices share an element.
        ec = colour[elems]
        for row in ec:
            row = row[row >= 0]
            assert len(row) == len(set(row.tolist()))


def test_serial_equals_group_major_and_thread_count(oracle):
    m = cube_mesh(6, "tet", 0.3, seed=9)
    results = []
    for order, threads in [(0, 1), (1, 1), (1, 4)]:
        o = oracle.OracleMesh.from_cube(m)
        o.smooth("gd", 2, threads=threads, order=order)
        results.append(o.positions())
    assert np.array_equal(results[0], results[1])
    assert np.array_equal(results[1], results[2])


def _single_free_vertex(oracle, kind, offset):
    m = cube_mesh(2, kind, 0.0, seed=1)
    pos = m.pos.copy()
    centre = int(np.argmin(np.sum((pos - 0.5) ** 2, axis=1)))
    pos[centre] += offset
    o = oracle.OracleMesh(pos, m.fixed, m.tets, m.hexs)
    return o


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_single_vertex_improves_towards_brute_force(oracle, smoother, kind):
    """One free vertex in a 2x2x2 cube: the smoother's patch quality never
    decreases and stays below the dense-grid optimum."""
    offset = np.array([0.11, -0.07, 0.05])
    o = _single_free_vertex(oracle, kind, offset)
    v = o.nverts - 1           # the only movable vertex is renumbered last
    q0 = o.patch_quality(v)
    start = o.positions()
    best = -np.inf
    grid = np.linspace(-0.2, 0.2, 21)
    for dx in grid:
        for dy in grid:
            for dz in grid:
                p = start.copy()
                p[v] += [dx, dy, dz]
                o.set_positions(p)
                best = max(best, o.patch_quality(v))
    o.set_positions(start)
    o.smooth(smoother, 3, threads=1, spawn_coeff=0.25)
    q1 = o.patch_quality(v)
    assert q1 >= q0
    assert q1 <= best + 1e-3
    assert q1 > q0 + 0.25 * (best - q0)


def test_gd_fixed_point_on_regular_hex_grid(oracle):
    m = cube_mesh(4, "hex", 0.0, seed=1)
    o = oracle.OracleMesh.from_cube(m)
    before = o.positions()
    o.smooth("gd", 1, threads=1)
    assert np.max(np.abs(o.positions() - before)) < 1e-12


def test_libm_modes_agree_closely(oracle):
    """glibc pow vs the project's pow: the C1-style run ends within a tiny distance."""
    m = cube_mesh(8, "tet", 0.3, seed=1)
    out = []
    for mode in (0, 1):
        oracle.set_libm(mode)
        o = oracle.OracleMesh.from_cube(m)
        r = o.smooth("nm", 3, threads=4)
        out.append((o.positions(), r))
    oracle.set_libm(1)
    d = np.max(np.abs(out[0][0] - out[1][0]))
    assert d < 1e-3 * (1.0 / 8)
    assert abs(out[0][1]["mean"][-1] - out[1][1]["mean"][-1]) < 1e-6


@pytest.mark.parametrize("smoother", ["gd", "getme"])
def test_autopilot_schedule(oracle, smoother):
    """The autopilot (AbstractSmoother.cpp:186-197): the sweeps are those of
    the fixed-count run, and the run stops at the first record k >= 1 whose
    harmonic-mean gain is at most min_qual_threshold and whose minimum gain is
    at most qual_mean_threshold.  The thresholds differ (a gain above one but
    not the other), so comparing either gain with the other threshold moves
    the stop."""
    m = cube_mesh(6, "tet", 0.35, seed=4)
    fixed = oracle.OracleMesh.from_cube(m)
    ref = fixed.smooth(smoother, 12, threads=1)
    t_mean, t_min = 2e-4, 5e-3
    auto = oracle.OracleMesh.from_cube(m)
    r = auto.smooth_autopilot(smoother, 12, min_qual_threshold=t_mean, qual_mean_threshold=t_min, threads=1)
    k = r["passes"]
    assert 1 <= k < 12
    assert np.array_equal(r["min"], ref["min"][:k + 1])
    assert np.array_equal(r["mean"], ref["mean"][:k + 1])

    def goes_on(j):
        return (ref["mean"][j] - ref["mean"][j - 1] > t_mean) or (ref["min"][j] - ref["min"][j - 1] > t_min)

    assert all(goes_on(j) for j in range(1, k))
    assert not goes_on(k)
    # The stop depends on which gain meets which threshold.
    swapped = oracle.OracleMesh.from_cube(m).smooth_autopilot(smoother, 12, min_qual_threshold=t_min,
                                                              qual_mean_threshold=t_mean, threads=1)
    first = next((j for j in range(1, 13) if not ((ref["mean"][j] - ref["mean"][j - 1] > t_min) or
                                                   (ref["min"][j] - ref["min"][j - 1] > t_mean))), 12)
    assert swapped["passes"] == first
    assert first != k
    # Positions after k autopilot sweeps are those after k fixed sweeps.
    again = oracle.OracleMesh.from_cube(m)
    again.smooth(smoother, k, threads=1)
    assert np.array_equal(again.positions(), auto.positions())


def test_autopilot_limits(oracle):
    m = cube_mesh(4, "tet", 0.3, seed=2)
    o = oracle.OracleMesh.from_cube(m)
    assert o.smooth_autopilot("gd", 0)["passes"] == 0          # no sweep after record 0
    o = oracle.OracleMesh.from_cube(m)
    assert o.smooth_autopilot("gd", 3, min_qual_threshold=-1.0)["passes"] == 3   # never stops: the cap
    o = oracle.OracleMesh.from_cube(m)
    assert o.smooth_autopilot("gd", 5, min_qual_threshold=1.0, qual_mean_threshold=1.0)["passes"] == 1


def test_metric_wise_zero_length_edge_ends():
    """The metric-wise measurer's adaptive walk on a zero-length edge (two
    coincident vertices): the pair difference is NaN, which never passes the
    threshold; the walk must end (NaN quality for the elements holding the
    edge) instead of subdividing without bound.  Runs in a child process
    under a time limit so that a regression cannot stall the suite."""
    import subprocess
    import sys
    import textwrap
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = textwrap.dedent("""
        import numpy as np
        from oracle import pyoracle as po
        from gpumesh_amd import cube_mesh
        m = cube_mesh(2, "tet", 0.0, seed=1)
        inner = int(np.where(m.fixed == 0)[0][0])
        t = m.tets[np.any(m.tets == inner, axis=1)][0]
        pos = m.pos.copy()
        pos[inner] = pos[int(t[t != inner][0])]
        om = po.OracleMesh(pos, m.fixed, m.tets, m.hexs)
        om.set_crew("mean_ratio", "analytic", scaling=2.0)
        q = om.element_qualities()
        print(int(np.isnan(q).sum()), int(np.isfinite(q).sum()))
    """)
    r = subprocess.run([sys.executable, "-c", code], cwd=root, timeout=120, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["6", "42"]
