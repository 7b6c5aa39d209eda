"""Offline rendering of the octree and the particles to PNG (no window, no GUI toolkit).

The picture is the one the reference's interactive viewer draws (``gui.py:12-70``):
exactly-ACTIVATED cells of every level in light blue with a darker one-pixel border,
exactly-GHOST cells in blue, and a 3x3 dot on every finest-level node that is
activated-not-T (red), a T-junction (green) or ghost-only (dark blue); particles are disks of
radius 0.003 (in domain units) in their ``c_p`` colour.  Finer levels are drawn over coarser
ones, as the reference's level loop does.  For a 3D simulation the same picture is drawn for
one slice through the domain, the counterpart of ``gui_3d.py:34-67``'s cell outlines.

Everything here works on plain arrays, so the images of the GPU simulator and of any other
source of masks can be compared pixel for pixel.
"""
import struct
import zlib

import numpy as np

ACTIVATED, GHOST, T_JUNCTION = 1, 2, 4

# colours of gui.py:21-42
ACTIVE_FILL, ACTIVE_EDGE = (0.28, 0.68, 0.99), (0.18, 0.58, 0.88)
GHOST_FILL, GHOST_EDGE = (0.0, 0.0, 1.0), (0.0, 0.0, 0.77)
NODE_ACTIVE, NODE_T, NODE_GHOST = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 0.35)
PARTICLE_RADIUS = 0.003


def mask_image(cell_masks, node_mask, coarsest_size, res=768, show_ghost=True, show_node=True):
    """RGB float image ``img[i, j]`` (i along axis 0 of the masks, j along axis 1) of 2D
    cell masks (``cell_masks[l]`` covering at least the level's ``coarsest_size * 2**l``
    cells) and the finest-level node mask; ``gui.py:14-42``.  A pixel maps to cell ``i // scale`` with
    ``scale = res // extent``; pixels past the last cell stay black."""
    L = len(cell_masks)
    img = np.zeros((res, res, 3), np.float32)
    pix = np.arange(res)
    for l in range(L):
        n = coarsest_size * 2 ** l
        s = res // n
        if s == 0:
            raise ValueError(f"res {res} is below the level-{l} extent {n}")
        m = np.asarray(cell_masks[l])
        if m.ndim != 2 or min(m.shape) < n:
            raise ValueError(f"cell_masks[{l}] has shape {m.shape}, expected at least {(n, n)}")
        m = m[:n, :n]  # fields are allocated padded past the domain (utils.py:44-45)
        ci = pix // s
        inside = ci < n
        mm = m[np.ix_(np.minimum(ci, n - 1), np.minimum(ci, n - 1))]
        valid = inside[:, None] & inside[None, :]
        e1 = (pix % s == 0) | (pix % s == s - 1)
        edge = e1[:, None] | e1[None, :]
        act = valid & (mm == ACTIVATED)
        img[act & ~edge] = ACTIVE_FILL
        img[act & edge] = ACTIVE_EDGE
        if show_ghost:
            gh = valid & (mm == GHOST)
            img[gh & ~edge] = GHOST_FILL
            img[gh & edge] = GHOST_EDGE
    if show_node:
        fs = coarsest_size * 2 ** (L - 1)
        nm = np.asarray(node_mask)
        fscale = res // fs
        ni = np.arange(0, res, fscale)
        ni = ni[ni // fscale < nm.shape[0]]
        k = ni // fscale
        v = nm[np.ix_(k, k)]
        colour = np.full(v.shape, -1)
        colour[(v & GHOST != 0) & (v & (ACTIVATED | T_JUNCTION) == 0)] = 2
        colour[v & T_JUNCTION != 0] = 1
        colour[(v & ACTIVATED != 0) & (v & T_JUNCTION == 0)] = 0
        pal = np.array([NODE_ACTIVE, NODE_T, NODE_GHOST], np.float32)
        a, b = np.nonzero(colour >= 0)
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                pi, pj = ni[a] + di, ni[b] + dj
                ok = (pi >= 0) & (pi < res) & (pj >= 0) & (pj < res)
                img[pi[ok], pj[ok]] = pal[colour[a[ok], b[ok]]]
    return img


def draw_particles(img, x, colours=None, radius=PARTICLE_RADIUS):
    """Disks of ``radius`` (domain units) at the 2D positions ``x`` (domain [0, 1)^2) drawn
    into ``img[i, j]`` in place; ``colours`` (n, 3) in [0, 1], white when omitted."""
    res = img.shape[0]
    x = np.asarray(x, np.float64).reshape(-1, 2)
    col = np.ones((x.shape[0], 3), np.float32) if colours is None else np.asarray(colours, np.float32).reshape(-1, 3)
    r = radius * res
    cx, cy = x[:, 0] * res, x[:, 1] * res
    R = int(np.ceil(r))
    for di in range(-R, R + 1):
        for dj in range(-R, R + 1):
            pi = np.floor(cx).astype(np.int64) + di
            pj = np.floor(cy).astype(np.int64) + dj
            d2 = (pi + 0.5 - cx) ** 2 + (pj + 0.5 - cy) ** 2
            ok = (d2 <= r * r) & (pi >= 0) & (pi < res) & (pj >= 0) & (pj < res)
            img[pi[ok], pj[ok]] = col[ok]
    return img


def to_rgb8(img):
    """``img[i, j]`` (i rightwards, j upwards) to an 8-bit (rows, cols, 3) picture, top row first."""
    return np.ascontiguousarray(np.clip(np.rint(img * 255.0), 0, 255).astype(np.uint8).transpose(1, 0, 2)[::-1])


def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def write_png(path, rgb8):
    """8-bit RGB PNG (colour type 2, no interlace, filter 0 on every row)."""
    rgb8 = np.ascontiguousarray(rgb8, np.uint8)
    h, w, c = rgb8.shape
    if c != 3:
        raise ValueError("write_png takes (rows, cols, 3) uint8")
    raw = np.concatenate([np.zeros((h, 1), np.uint8), rgb8.reshape(h, w * 3)], axis=1).tobytes()
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)))
        f.write(_chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(_chunk(b"IEND", b""))


def read_png(path):
    """Reads back what ``write_png`` writes (8-bit RGB, filter 0); checks every chunk's CRC."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        if zlib.crc32(kind + body) & 0xFFFFFFFF != crc:
            raise ValueError(f"bad CRC in {kind!r}")
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[:4]
    if depth != 8 or ctype != 2:
        raise ValueError("only 8-bit RGB is read")
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    if raw[:, 0].any():
        raise ValueError("only filter type 0 is read")
    return raw[:, 1:].reshape(h, w, 3).copy()


def slice_masks(cell_masks, node_mask, coarsest_size, axis=2, depth=0.5):
    """2D cell and node masks of the plane ``x[axis] = depth`` of a 3D octree: at level l the
    cells ``int(depth * n_l)`` along ``axis``, and the finest nodes nearest to the plane."""
    L = len(cell_masks)
    cells = []
    for l in range(L):
        n = coarsest_size * 2 ** l
        k = min(int(depth * n), n - 1)
        cells.append(np.take(np.asarray(cell_masks[l]), k, axis=axis))
    fs = coarsest_size * 2 ** (L - 1)
    nm = np.asarray(node_mask)
    k = min(int(depth * fs + 0.5), nm.shape[axis] - 1)
    return cells, np.take(nm, k, axis=axis)


def render(sim, path=None, res=768, show_particles=True, show_ghost=True, show_node=True, axis=2, depth=0.5):
    """Picture of a simulator's current octree and particles (``gui.py:64-70``); for 3D, of
    the slice ``x[axis] = depth`` with the particles within half a finest cell of it.
    Writes a PNG when ``path`` is given; returns the 8-bit (rows, cols, 3) picture."""
    cells = [sim.cell_mask[l].to_numpy() for l in range(sim.level)]
    nodes = sim.node_mask.to_numpy()
    x = sim.x_p.to_numpy()
    col = sim.c_p.to_numpy()
    if sim.dim == 3:
        cells, nodes = slice_masks(cells, nodes, sim.coarsest_size, axis, depth)
        keep = [a for a in range(3) if a != axis]
        near = np.abs(x[:, axis] - depth) <= 0.5 / sim.finest_size
        x, col = x[near][:, keep], col[near]
    img = mask_image(cells, nodes, sim.coarsest_size, res, show_ghost, show_node)
    if show_particles:
        if not np.any(col):
            col = None
        draw_particles(img, x, col)
    rgb = to_rgb8(img)
    if path is not None:
        write_png(path, rgb)
    return rgb
