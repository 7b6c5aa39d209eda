"""Offline renderer (gui.py:12-70 picture as a PNG): the PNG file format, colours of a
hand-made octree at known pixels, and, on an octree built by the oracle, the area of the
ACTIVATED cells and the node-dot count, each counted independently of the renderer."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def R(pkg):
    return pkg.render


def rgb(c):
    return tuple(int(round(255 * v)) for v in c)


def at(rgb8, i, j):
    """Colour of img[i, j] (i rightwards, j upwards) in the top-row-first picture."""
    return tuple(int(v) for v in rgb8[rgb8.shape[0] - 1 - j, i])


def test_png_round_trip_and_crc(R, tmp_path):
    img = np.random.default_rng(0).integers(0, 256, (37, 53, 3), dtype=np.uint8)
    p = tmp_path / "a.png"
    R.write_png(p, img)
    data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n" and data[12:16] == b"IHDR"
    assert np.array_equal(R.read_png(p), img)
    bad = bytearray(data)
    bad[40] ^= 0xFF  # inside the IDAT body
    p.write_bytes(bytes(bad))
    with pytest.raises(ValueError):
        R.read_png(p)


def test_hand_made_octree_colours(R):
    c, res = 2, 16  # level 0: 2x2 cells of 8 px; level 1: 4x4 cells of 4 px; nodes every 4 px
    l0 = np.zeros((2, 2), np.int32)
    l1 = np.zeros((4, 4), np.int32)
    l0[0, 0] = R.ACTIVATED
    l1[2, 0] = R.ACTIVATED
    l1[3, 3] = R.GHOST
    l0[1, 1] = R.GHOST | R.ACTIVATED  # neither exactly A nor exactly G: not drawn
    nm = np.zeros((5, 5), np.int32)
    nm[1, 1] = R.ACTIVATED
    nm[2, 2] = R.ACTIVATED | R.T_JUNCTION
    nm[3, 3] = R.GHOST
    nm[0, 3] = R.ACTIVATED | R.GHOST
    nm[3, 1] = R.GHOST | R.T_JUNCTION
    img = R.to_rgb8(R.mask_image([l0, l1], nm, c, res, show_node=False))
    assert at(img, 3, 2) == rgb(R.ACTIVE_FILL) and at(img, 0, 2) == rgb(R.ACTIVE_EDGE) and at(img, 7, 7) == rgb(R.ACTIVE_EDGE)
    assert at(img, 9, 1) == rgb(R.ACTIVE_FILL) and at(img, 8, 1) == rgb(R.ACTIVE_EDGE)
    assert at(img, 13, 13) == rgb(R.GHOST_FILL) and at(img, 15, 12) == rgb(R.GHOST_EDGE)
    assert at(img, 10, 10) == (0, 0, 0) and at(img, 1, 12) == (0, 0, 0)
    img = R.to_rgb8(R.mask_image([l0, l1], nm, c, res, show_ghost=False, show_node=False))
    assert at(img, 13, 13) == (0, 0, 0)
    img = R.to_rgb8(R.mask_image([l0, l1], nm, c, res))
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            assert at(img, 4 + di, 4 + dj) == rgb(R.NODE_ACTIVE)
            assert at(img, 8 + di, 8 + dj) == rgb(R.NODE_T)
            assert at(img, 12 + di, 12 + dj) == rgb(R.NODE_GHOST)
            assert at(img, 12 + di, 4 + dj) == rgb(R.NODE_T)
    assert at(img, 0, 12) == rgb(R.NODE_ACTIVE) and at(img, 1, 13) == rgb(R.NODE_ACTIVE)  # clipped at the edge


def test_oracle_octree_areas_and_node_dots(R, pkg, O):
    n, L, c, res = 20000, 4, 16, 768
    st = {}

    class Rec:
        n_particles, p_mass = n, 1.0

        def __getattr__(self, k):
            class P:
                def from_numpy(self, a):
                    st[k] = np.asarray(a)
            return P()

    pkg.scenes.reference_case0_2d(0)(Rec())
    ora = O.OracleSim(2, L, c, n, fp64=False)
    ora.set_particles(st["x_p"], st["v_p"], st["F_p"], np.zeros((n, 2, 2)), st["m_p"], st["g_p"])
    ora.refine()
    cells = [ora.get_level(l)[0] for l in range(L)]
    nm = ora.get_node_mask()
    img = R.to_rgb8(R.mask_image(cells, nm, c, res, show_ghost=False, show_node=False))
    flat = img.reshape(-1, 3)
    is_col = lambda col: np.all(flat == np.array(rgb(col)), axis=1)
    area = sum(int((cells[l] == R.ACTIVATED).sum()) * (res // (c * 2 ** l)) ** 2 for l in range(L))
    assert area > 0 and int((is_col(R.ACTIVE_FILL) | is_col(R.ACTIVE_EDGE)).sum()) == area
    img = R.to_rgb8(R.mask_image(cells, nm, c, res))
    flat = img.reshape(-1, 3)
    fscale = res // (c * 2 ** (L - 1))
    i = np.arange(nm.shape[0])
    span = np.where(i * fscale < res, np.minimum(i * fscale + 1, res - 1) - np.maximum(i * fscale - 1, 0) + 1, 0)
    w = span[:, None] * span[None, :]
    for col, sel in ((R.NODE_ACTIVE, (nm & (R.ACTIVATED | R.T_JUNCTION)) == R.ACTIVATED),
                     (R.NODE_T, (nm & R.T_JUNCTION) != 0),
                     (R.NODE_GHOST, (nm & (R.ACTIVATED | R.GHOST | R.T_JUNCTION)) == R.GHOST)):
        assert sel.any()
        assert int(is_col(col).sum()) == int(w[sel].sum()), col


def test_particles_orientation_and_slice(R):
    img = np.zeros((100, 100, 3), np.float32)
    R.draw_particles(img, np.array([[0.105, 0.905]]), np.array([[1.0, 0.5, 0.0]]), radius=0.02)
    rgb8 = R.to_rgb8(img)
    rows, cols = np.nonzero(rgb8[..., 0])
    assert abs(rows.mean() - 9.5) < 0.6 and abs(cols.mean() - 10.5) < 0.6  # top-left of the picture
    assert rows.size == pytest.approx(np.pi * 2.0 ** 2, abs=4)
    c = 4
    cells = [np.arange(64).reshape(4, 4, 4), np.arange(512).reshape(8, 8, 8)]
    nm = np.arange(9 ** 3).reshape(9, 9, 9)
    s, n2 = R.slice_masks(cells, nm, c, axis=1, depth=0.3)
    assert np.array_equal(s[0], cells[0][:, 1, :]) and np.array_equal(s[1], cells[1][:, 2, :])
    assert np.array_equal(n2, nm[:, 2, :])
