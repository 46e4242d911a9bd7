"""Batches: many resident meshes smoothed by fused launches.

A batch member's kernels, launch order within the member and reduction tree
are the ones it gets alone, so every member must come out bit-identical to
the same mesh smoothed alone -- positions and every pass record, in FP64 and
in the mixed-precision modes -- eagerly and when one pass body is replayed
from a HIP graph.  The FP64 single-mesh path is itself pinned bit-exact to the
oracle by test_gpu_parity.py, so FP64 batch results are pinned to the oracle
transitively.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuBatch, GpuMesh, cube_mesh

pytestmark = pytest.mark.gpu

COEFF = {"nm": 1.0 / 20.0, "gd": 1.0 / 20.0, "spawn": 0.25, "getme": 1.0 / 20.0}


def _mixed_mesh():
    t = cube_mesh(4, "tet", 0.0, seed=1)
    h = cube_mesh(4, "hex", 0.0, seed=1)
    cx_t = t.pos[t.tets].mean(axis=1)[:, 0]
    cx_h = h.pos[h.hexs].mean(axis=1)[:, 0]
    rng = np.random.default_rng(2)
    pos = t.pos + np.where(t.fixed[:, None] == 1, 0.0, rng.uniform(-0.05, 0.05, t.pos.shape))
    return pos, t.fixed, t.tets[cx_t > 0.5], h.hexs[cx_h < 0.5]


def _members():
    out = []
    for n, kind, seed in ((6, "tet", 3), (9, "tet", 4), (7, "hex", 5), (5, "hex", 6), (3, "tet", 7)):
        m = cube_mesh(n, kind, 0.3, seed=seed)
        out.append((m.pos, m.fixed, m.tets, m.hexs))
    out.append(_mixed_mesh())
    return out


def _make(members, crew=None):
    ms = [GpuMesh(*a) for a in members]
    if crew:
        for m in ms:
            m.set_crew(**crew)
    return ms


def _alone(members, smoother, passes, precision, crew=None):
    res = []
    for m in _make(members, crew):
        r = m.smooth(smoother, passes, precision=precision, kernel="wave", spawn_coeff=COEFF[smoother])
        res.append((m.positions(), r))
    return res


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32_polish"])
@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_batch_members_match_alone(smoother, precision, graph):
    members = _members()
    passes = 3
    ref = _alone(members, smoother, passes, precision)
    ms = _make(members)
    b = GpuBatch(ms)
    r = b.smooth(smoother, passes, precision=precision, kernel="wave", spawn_coeff=COEFF[smoother], graph=graph)
    for i, m in enumerate(ms):
        pos, rr = ref[i]
        assert _same(m.positions(), pos), (i, smoother, precision, graph)
        assert _same(r["min"][i], rr["min"]), i
        assert _same(r["inv_sum"][i], rr["inv_sum"]), i
        assert np.array_equal(r["hist"][i], rr["hist"]), i
    b.close()


@pytest.mark.parametrize("smoother", ["nm", "spawn", "getme"])
def test_batch_anisotropic_crew_matches_alone(smoother):
    members = _members()[:4]
    crew = {"evaluator": "metric_conformity", "sampler": "analytic", "scaling": 2.0}
    ref = _alone(members, smoother, 2, "fp64", crew)
    ms = _make(members, crew)
    b = GpuBatch(ms)
    r = b.smooth(smoother, 2, precision="fp64", kernel="wave", spawn_coeff=COEFF[smoother])
    for i, m in enumerate(ms):
        assert _same(m.positions(), ref[i][0]), i
        assert _same(r["inv_sum"][i], ref[i][1]["inv_sum"]), i


@pytest.mark.parametrize("precision", ["fp64", "fp32_polish"])
@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_batch_texture_crew_matches_alone(smoother, precision):
    """Members sharing one texture grid (metric conformity, the shell field
    sampled on 12^3 cells) come out as they do alone, from graph replay."""
    from gpumesh_amd.samplers import shell_texture
    members = _members()[:4]
    tex = shell_texture(members[0][0], 12, 3.0)
    crew = {"evaluator": "metric_conformity", "sampler": "texture", "scaling": 3.0, "texture": tex}
    ref = _alone(members, smoother, 2, precision, crew)
    ms = _make(members, crew)
    b = GpuBatch(ms)
    r = b.smooth(smoother, 2, precision=precision, kernel="wave", spawn_coeff=COEFF[smoother], graph=True)
    for i, m in enumerate(ms):
        assert _same(m.positions(), ref[i][0]), i
        assert _same(r["min"][i], ref[i][1]["min"]), i
        assert _same(r["inv_sum"][i], ref[i][1]["inv_sum"]), i
    b.close()


def test_batch_refuses_members_with_different_textures():
    from gpumesh_amd import GpuMeshError
    from gpumesh_amd.samplers import shell_texture
    members = _members()[:2]
    ms = _make(members)
    ms[0].set_crew("metric_conformity", "texture", 3.0, shell_texture(members[0][0], 12, 3.0))
    ms[1].set_crew("metric_conformity", "texture", 3.0, shell_texture(members[0][0], 12, 2.0))
    with pytest.raises(GpuMeshError):
        GpuBatch(ms)
    ms[1].set_crew("metric_conformity", "texture", 3.0, shell_texture(members[0][0], 12, 3.0))
    GpuBatch(ms).close()


def test_graph_replays_match_eager_with_work_between():
    """The captured pass bodies are replayed in several runs, with eager work
    of other handles in between; every replay must reproduce the eager run."""
    members = _members()
    eager_ms = _make(members)
    eb = GpuBatch(eager_ms)
    er = eb.smooth("gd", 4, precision="fp32_polish", kernel="wave")
    eager_pos = [m.positions() for m in eager_ms]
    ms = _make(members)
    b = GpuBatch(ms)
    other = GpuMesh(*members[1])
    for rep in range(3):
        # Reset to the starting positions (renumbered order).
        for m, a in zip(ms, members):
            perm = m.permutation()
            pos = np.empty_like(a[0])
            pos[perm] = a[0]
            m.set_positions(pos)
        r = b.smooth("gd", 4, precision="fp32_polish", kernel="wave", graph=True)
        for i, m in enumerate(ms):
            assert _same(m.positions(), eager_pos[i]), (rep, i)
            assert _same(r["inv_sum"][i], er["inv_sum"][i]), (rep, i)
        other.smooth("nm", 1, precision="fp64", kernel="plain")
        other.evaluate()


def test_autotune_keeps_positions_and_results():
    members = _members()[:4]
    ms = _make(members)
    b = GpuBatch(ms)
    before = [m.positions() for m in ms]
    (ppg, tile), times = b.autotune("spawn", precision="fp32", passes=1)
    assert ppg in (1, 2, 4, 8) and tile in (64, 128, 256)
    assert times.shape == (4, 3)
    assert times[[1, 2, 4, 8].index(ppg), [64, 128, 256].index(tile)] == times[times >= 0].min()
    for m, p in zip(ms, before):
        assert _same(m.positions(), p)
    # The tuned shape changes the launch, not the arithmetic.
    r = b.smooth("spawn", 2, precision="fp64", kernel="wave", spawn_coeff=0.25)
    ref = _alone(members, "spawn", 2, "fp64")
    for i, m in enumerate(ms):
        assert _same(m.positions(), ref[i][0]), i
        assert _same(r["inv_sum"][i], ref[i][1]["inv_sum"]), i


@pytest.mark.parametrize("smoother", ["gd", "spawn"])
def test_every_launch_shape_gives_the_same_bits(smoother):
    """Every shape the autotuner accepts runs the candidate-parallel kernel on
    every single-kind list, with the same arithmetic as the default shape:
    FP32 results are bit-identical to the members smoothed alone."""
    members = _members()[:3]
    ref = _alone(members, smoother, 2, "fp32")
    probe = GpuBatch(_make(members))
    _, times = probe.autotune(smoother, precision="fp32", passes=1)
    assert (times >= 0).any()
    for a, ppg in enumerate((1, 2, 4, 8)):
        for t, tile in enumerate((64, 128, 256)):
            if times[a, t] < 0:
                continue
            ms = _make(members)
            b = GpuBatch(ms)
            b.set_shape(smoother, ppg, tile)
            b.smooth(smoother, 2, precision="fp32", kernel="wave", spawn_coeff=COEFF[smoother])
            for i, m in enumerate(ms):
                assert _same(m.positions(), ref[i][0]), (ppg, tile, i)
