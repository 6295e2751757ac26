"""Host side of cbmv_amd.main (no GPU): image reading as 8-bit greyscale
(scipy.misc.imread(..., mode='L') of test.py:506-507 is PIL's 'L'
conversion), the cache file name of main.py:55, and that a cache holding
Python objects is refused (numpy.load without pickle)."""
import numpy as np
import pytest
from PIL import Image

from cbmv_amd import main as cmain


def test_read_grey_png_and_npy(tmp_path):
    rng = np.random.default_rng(0)
    rgb = rng.integers(0, 256, (7, 9, 3)).astype(np.uint8)
    Image.fromarray(rgb).save(tmp_path / "c.png")
    g = cmain.read_grey(str(tmp_path / "c.png"))
    r, gg, b = (rgb[..., k].astype(np.int64) for k in range(3))
    # ITU-R 601-2 luma with PIL's rounding: (R*19595 + G*38470 + B*7471 + 0x8000) >> 16
    np.testing.assert_array_equal(g, ((r * 19595 + gg * 38470 + b * 7471 + 0x8000) >> 16).astype(np.uint8))
    np.save(tmp_path / "g.npy", g)
    np.testing.assert_array_equal(cmain.read_grey(str(tmp_path / "g.npy")), g)


def test_cost_file_name_and_proba_layout():
    assert cmain.cost_file("./results", "Adirondack") == "./resultsAdirondack.prob.npy"
    p = cmain.proba_from_cost(np.array([[[0.25, 1.0]]], np.float32))
    np.testing.assert_array_equal(p, [[0.25, 0.75], [1.0, 0.0]])


def test_cache_loader_refuses_objects_and_wrong_shapes(tmp_path):
    f = tmp_path / "x.prob.npy"
    np.save(f, np.array([{"a": 1}], dtype=object), allow_pickle=True)
    with pytest.raises(ValueError):
        cmain.load_proba(str(f), 1)
    np.save(f, np.zeros((6, 2)))
    assert cmain.load_proba(str(f), 6).shape == (6, 2)
    with pytest.raises(ValueError, match="expected"):
        cmain.load_proba(str(f), 7)
