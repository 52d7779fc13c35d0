"""rad_cast_static (the allocation-free variant of rad_cast) through the C
ABI: on two calls in a row, the second with fewer rays, its hits, leaf counts
and test counts equal rad_cast's on the same rays; sizes beyond its limit are
refused with RAD_EINVAL before any device work."""
import numpy as np
import pytest

import radgpu
from radgpu import scenes

pytestmark = pytest.mark.gpu


def _rays(rec, n, seed):
    rng = np.random.default_rng(seed)
    live = np.nonzero(rec["parent"] >= 0)[0]
    src, dst, v = rng.choice(live, n), rng.choice(live, n), rng.integers(0, 3, n)
    rays = np.zeros(n, radgpu.RAY_DTYPE)
    rays["p"] = rec["p"][dst, v]
    rays["v"] = rec["c"][src] - rec["p"][dst, v]
    rays["ban"] = rec["parent"][dst]
    rays["target"] = rec["parent"][src]
    return rays


def test_cast_static_equals_cast_on_two_calls():
    s = scenes.hall_scene(2, n_boxes=60)
    print("cast_static on hall with 60 boxes, %d triangles" % s.n, flush=True)
    with radgpu.Engine("-converge", "1200", "-triangle", "1000000", "-quiet", "1") as eng:
        eng.set_scene(s).prepare_host().upload()
        rec, _ = eng.records()
        for n, seed in ((radgpu.CAST_STATIC_MAX, 1), (5000, 2)):
            rays = _rays(rec, n, seed)
            want = eng.cast(rays)
            got = eng.cast_static(rays)
            for w, g in zip(want, got):
                assert np.array_equal(w, g), n
            assert (want[0] >= 0).any() and (want[0] == -1).any()
        too_many = np.zeros(radgpu.CAST_STATIC_MAX + 1, radgpu.RAY_DTYPE)
        with pytest.raises(radgpu.RadError) as ei:
            eng.cast_static(too_many)
        assert ei.value.status == radgpu.RAD_EINVAL and "limit" in str(ei.value)
