"""Host-side checks of SceneBatch that need no GPU: scenes whose step the
grouped kernels do not cover are refused before anything is allocated."""
import pytest

from conftest import load_scene


@pytest.mark.parametrize("flag", ["fbm", "wcsph"])
def test_scene_batch_refuses_unbatched_blocks(flag):
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.batch import SceneBatch
    cb, sb = load_scene("dambreak_2d_8k")
    sim = simulation_from_dicts(cb, sb, device="cpu", **{flag: True})
    with pytest.raises(ValueError, match="FBM" if flag == "fbm" else "WCSPH"):
        SceneBatch([sim])


def test_scene_batch_refuses_empty_and_mixed_precision():
    from tichi_sph_amd import simulation_from_dicts
    from tichi_sph_amd.batch import SceneBatch
    cb, sb = load_scene("dambreak_2d_8k")
    with pytest.raises(ValueError):
        SceneBatch([])
    a = simulation_from_dicts(cb, sb, device="cpu", precision=4)
    b = simulation_from_dicts(cb, sb, device="cpu", precision=8)
    with pytest.raises(ValueError, match="precision"):
        SceneBatch([a, b])
