"""Host-side pieces of the learned warm start that need no GPU."""
import dataclasses

import numpy as np


def test_seeded_case_on_a_shared_mesh_equals_a_rebuilt_one():
    """grf_case with `like` reuses the mesh arrays of another model of the same grid and gives the
    same case, field for field, as building the mesh again."""
    from actfemtide_amd import surrogate as S
    base = S.grf_case((6, 5, 8), 0)
    a = S.grf_case((6, 5, 8), 7)
    b = S.grf_case((6, 5, 8), 7, like=base)
    for f in dataclasses.fields(a):
        x, y = getattr(a, f.name), getattr(b, f.name)
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y, equal_nan=True), f.name
        else:
            assert x == y, f.name
    assert not np.array_equal(a.sigma_crust, base.sigma_crust)
