"""The FP32 optimisers' decisions against the FP64 oracle on the 1M-vertex
Kuhn cube (100^3 cells, jitter 0.35 h, seed 2), one colour step at a time.

Positions enter an FP32 pass rounded to FP32 and every patch quality is
evaluated on FP32 affine corner terms, so an FP32 choice can differ from the
FP64 one only where the FP64 candidates are close.  For 1500 sampled
vertices of colour groups 0 and 5 (starting from the state the GPU reached
before the group):
  * spawn search: where in both sweeps the oracle's best lattice point beats
    its runner-up by more than GAP, the FP32 result is the oracle's point;
    over all vertices the FP64 patch quality at the FP32 choice is within
    QREL of the oracle's best;
  * gradient descent: where every line search of the oracle's run is decided
    by more than GAP, the final position is within 1e-3 L (L the local
    size) of the oracle's; over all vertices the patch quality is within QREL.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuMesh, cube_mesh

pytestmark = pytest.mark.gpu

N_SAMPLE = 1500
GAP = 1e-4        # FP32 inputs and affine terms move patch qualities by ~1e-5
QREL = 2e-5


@pytest.fixture(scope="module")
def c2(oracle):
    m = cube_mesh(100, "tet", 0.35, seed=2)
    gm = GpuMesh.from_cube(m)
    om = oracle.OracleMesh.from_cube(m)
    return m, gm, om, gm.positions()


def _lattice():
    off = [np.array([i, j, k], float) - 1.5 for k in range(4) for j in range(4) for i in range(4)]
    off[0] = np.zeros(3)
    return np.array(off)


def _step(gm, smoother, coeff, start, group):
    gm.set_positions(start)
    if group > 0:
        gm.smooth_groups(smoother, 0, group, "fp64", "wave", coeff)
    before = gm.positions()
    gm.smooth_groups(smoother, group, group + 1, "fp32", "wave", coeff)
    return before, gm.positions()


def _sample(gm, group):
    off, nodes = gm.groups()
    members = nodes[off[group]:off[group + 1]]
    return np.sort(np.random.default_rng(200 + group).choice(members, size=N_SAMPLE, replace=False))


@pytest.mark.parametrize("group", [0, 5])
def test_spawn_fp32_choices(c2, group):
    m, gm, om, start = c2
    coeff = 0.25
    before, after = _step(gm, "spawn", coeff, start, group)
    lat = _lattice()
    sample = _sample(gm, group)
    # Vertices of one colour group share no element: tracing them one after
    # another on the same state gives each its own colour-step result.
    om.set_positions(before)
    decided = 0
    q_best = {}
    for v in sample:
        v = int(v)
        s = coeff * om.local_size(v)
        tr = om.trace_vertex("spawn", v, spawn_coeff=coeff)
        q1, q2 = tr[:64, 3], tr[64:, 3]
        o1, o2 = int(np.argmax(q1)), int(np.argmax(q2))
        gap1 = q1[o1] - np.max(np.delete(q1, o1))
        gap2 = q2[o2] - np.max(np.delete(q2, o2))
        q_best[v] = q2[o2]
        # Where both sweeps are decided, the FP32 result is the oracle's
        # lattice point up to FP32 rounding (other points lie s/3 or more
        # away on some axis).
        if gap1 > GAP and gap2 > GAP:
            decided += 1
            ref = before[v] + lat[o1] * s + lat[o2] * (s / 3.0)
            assert np.max(np.abs(after[v] - ref)) <= 1e-3 * s, (v, o1, o2, after[v] - ref, s)
    om.set_positions(after)
    worst_rel = max((q_best[int(v)] - om.patch_quality(int(v))) / abs(q_best[int(v)]) for v in sample)
    print(f"spawn group {group}: {decided} of {N_SAMPLE} vertices with both sweeps decided by > {GAP}, all agree; "
          f"worst relative quality loss {worst_rel:.2e}")
    assert decided > N_SAMPLE // 4
    assert worst_rel <= QREL


@pytest.mark.parametrize("group", [0, 5])
def test_gradient_descent_fp32_moves(c2, group):
    m, gm, om, start = c2
    coeff = 0.05
    before, after = _step(gm, "gd", coeff, start, group)
    sample = _sample(gm, group)
    om.set_positions(before)
    decided = 0
    q_best, ref_pos, local = {}, {}, {}
    for v in sample:
        v = int(v)
        local[v] = om.local_size(v)
        q0 = om.patch_quality(v)
        tr = om.trace_vertex("gd", v)
        # Cycles of 6 gradient samples and 8 line points (a cycle that stops
        # on a zero gradient holds only its samples); the vertex ends on the
        # best line point of its last complete cycle.
        gaps, pos, q = [], before[v], q0
        k = 0
        while k + 14 <= tr.shape[0]:
            line = tr[k + 6:k + 14]
            b = int(np.argmax(line[:, 3]))
            gaps.append(line[b, 3] - np.max(np.delete(line[:, 3], b)))
            pos, q = line[b, :3], line[b, 3]
            k += 14
        q_best[v], ref_pos[v] = q, pos
        if gaps and min(gaps) > GAP:
            decided += 1
            assert np.max(np.abs(after[v] - pos)) <= 1e-3 * local[v], (v, after[v] - pos, local[v])
    om.set_positions(after)
    worst_rel = max((q_best[int(v)] - om.patch_quality(int(v))) / abs(q_best[int(v)]) for v in sample)
    print(f"gd group {group}: {decided} of {N_SAMPLE} vertices with every line search decided by > {GAP}; "
          f"worst relative quality loss {worst_rel:.2e}")
    assert decided > N_SAMPLE // 4
    assert worst_rel <= QREL
