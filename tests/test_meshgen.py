"""The native cube-mesh generator (gm_cube_mesh) against its numpy
restatement: identical coordinates, flags and connectivity, including cases
where the no-inversion repair halves the jitter several times."""
import numpy as np
import pytest

from gpumesh_amd.meshgen import _hex_min_corner_det, _tet_dets, cube_mesh, cube_mesh_numpy


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("n,kind,jitter,seed", [
    (1, "tet", 0.3, 1), (2, "hex", 0.4, 3), (5, "tet", 0.49, 7), (8, "hex", 0.49, 9), (20, "tet", 0.3, 1),
    (13, "hex", 0.45, 11), (17, "tet", 0.49, 12), (24, "tet", 0.35, 2), (11, "hex", 0.3, 1000),
    (9, "tet", 0.2, 2**64 - 1)])
def test_native_generator_matches_numpy(n, kind, jitter, seed):
    a = cube_mesh(n, kind, jitter, seed)
    b = cube_mesh_numpy(n, kind, jitter, seed)
    for f in ("pos", "fixed", "tets", "hexs"):
        assert np.array_equal(_bits(getattr(a, f)), _bits(getattr(b, f))), f


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_repair_leaves_no_thin_element(kind):
    n, frac = 10, 0.05
    m = cube_mesh(n, kind, 0.49, seed=5, min_det_fraction=frac)
    base = cube_mesh(n, kind, 0.0, seed=5)
    if kind == "tet":
        d = _tet_dets(m.pos, m.tets.astype(np.int64))
        ref = np.min(np.abs(_tet_dets(base.pos, base.tets[:6].astype(np.int64))))
    else:
        d = _hex_min_corner_det(m.pos, m.hexs.astype(np.int64))
        ref = (1.0 / n) ** 3
    assert d.min() >= frac * ref
    # The repair actually ran: some interior vertices kept less than full jitter.
    raw = cube_mesh(n, kind, 0.49, seed=5, min_det_fraction=-1e300)
    assert not np.array_equal(raw.pos, m.pos)
    assert np.array_equal(m.pos[m.fixed == 1], base.pos[base.fixed == 1])
