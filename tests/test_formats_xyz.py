"""The xyz trajectory writer keeps WriteConfig's layout (writeXYZFile,
src/WriteConfig.cpp:206-254): exact text of a small frame, and frames read
back."""
import numpy as np

from md_engine_amd import formats


def test_xyz_text_layout(tmp_path):
    p = tmp_path / "t.xyz"
    x = np.array([[0.5, 1.25, -3.0], [10.0, 0.000125, 2.5]])
    formats.write_xyz(p, x, [8, 1], [0.0, 0.0, 0.0], [12.5, 12.5, 12.5], append=False)
    assert p.read_text() == ("2\nbounds lo (0, 0, 0) hi (12.5, 12.5, 12.5)\n8 0.5 1.25 -3\n"
                             "1 10 0.000125 2.5\n")


def test_xyz_frames_round_trip(tmp_path):
    p = tmp_path / "t.xyz"
    rng = np.random.default_rng(0)
    for k in range(3):
        x = rng.uniform(0, 20, (5, 3))
        v = rng.normal(size=(5, 3))
        formats.write_xyz(p, x, [8, 1, 1, 0, 8], [0, 0, 0], [20, 20, 20], v=v)
    fr = formats.read_xyz(p)
    assert len(fr) == 3
    nums, vals, lo, hi = fr[-1]
    assert list(nums) == [8, 1, 1, 0, 8]
    assert vals.shape == (5, 6)
    assert np.allclose(vals[:, :3], x, rtol=1e-5) and np.allclose(vals[:, 3:], v, rtol=1e-5, atol=1e-5)
    assert np.allclose(hi, 20)
