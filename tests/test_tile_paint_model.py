"""CPU model of the index arithmetic of the tile paint (k_tile_key, k_tile_accumulate,
k_tile_boundary in pyrecon_amd/csrc/prh_cic.hip).

The model follows the kernels' integer arithmetic step by step (tile keys, the (PT + 1)^3 local array,
the split into cells only one tile reaches and boundary cells, the boundary pass over the lower
neighbours) with exact float sums, and checks that
- every index stays inside its buffer (tile ids < ntiles, local cells < (PT + 1)^3, mesh cells < n^3);
- no mesh cell is written by both passes (cells of empty tiles that no neighbour reaches are
  not written at all);
- the result equals a direct cloud-in-cell assignment (np.add.at) of the same weights,
for full and partial tiles, periodic wrap-around and meshes shorter than a tile.
"""
import itertools

import numpy as np
import pytest

PTB = 4                 # log2 of the tile edge (kPaintTileBits)
PT, PTH = 1 << PTB, (1 << PTB) + 1


def tile_grid(shape):
    return tuple((n + PT - 1) // PT for n in shape)


def tile_pos(t, shape, tiles):
    c = (t // (tiles[1] * tiles[2]), (t // tiles[2]) % tiles[1], t % tiles[2])
    o = tuple(PT * ci for ci in c)
    e = tuple(min(PT, n - oi) for n, oi in zip(shape, o))
    return c, o, e


def model_paint(shape, p, w):
    """p: (N, 3) mesh units in [0, n); w: (N,).  Returns the mesh and the write counts."""
    shape = tuple(shape)
    tiles = tile_grid(shape)
    ntiles = int(np.prod(tiles))
    b = np.floor(p).astype(int) % np.array(shape)
    d = p - b
    tile = ((b[:, 0] >> PTB) * tiles[1] + (b[:, 1] >> PTB)) * tiles[2] + (b[:, 2] >> PTB)
    assert tile.max() < ntiles and tile.min() >= 0
    key = (tile << (3 * PTB)) | (((b[:, 0] & (PT - 1)) * PT + (b[:, 1] & (PT - 1))) * PT + (b[:, 2] & (PT - 1)))
    assert key.max() < 2**32
    order = np.argsort(tile, kind='stable')
    mesh = np.zeros(shape)
    writes = np.zeros(shape, dtype=int)
    bnd = np.full((ntiles, PTH**3), np.nan)
    full = np.zeros(ntiles, dtype=bool)
    for t in range(ntiles):
        sel = order[tile[order] == t]
        if not len(sel): continue
        full[t] = True
        _, o, e = tile_pos(t, shape, tiles)
        acc = np.zeros(PTH**3)
        for i in sel:
            cell = key[i] & ((1 << (3 * PTB)) - 1)
            cx, cy, cz = cell >> (2 * PTB), (cell >> PTB) & (PT - 1), cell & (PT - 1)
            assert cx < e[0] and cy < e[1] and cz < e[2]
            for a, bb, c in itertools.product((0, 1), repeat=3):
                wt = (d[i, 0] if a else 1 - d[i, 0]) * (d[i, 1] if bb else 1 - d[i, 1]) * (d[i, 2] if c else 1 - d[i, 2])
                idx = ((cx + a) * PTH + (cy + bb)) * PTH + (cz + c)
                assert 0 <= idx < PTH**3
                acc[idx] += wt * w[i]
        for c in range(PTH**3):
            i, j, k = c // (PTH * PTH), (c // PTH) % PTH, c % PTH
            if i > e[0] or j > e[1] or k > e[2]: continue
            own = 1 <= i < e[0] and 1 <= j < e[1] and 1 <= k < e[2]
            if own:
                g = (o[0] + i, o[1] + j, o[2] + k)
                mesh[g] += acc[c]
                writes[g] += 1
            else:
                bnd[t, c] = acc[c]
    for t in range(ntiles):
        c, o, e = tile_pos(t, shape, tiles)
        em, nb = [], {}
        for dd in range(3):
            cm = tiles[dd] - 1 if c[dd] == 0 else c[dd] - 1
            em.append(min(PT, shape[dd] - PT * cm))
        for sub in range(8):
            cc = [(tiles[dd] - 1 if c[dd] == 0 else c[dd] - 1) if (sub >> (2 - dd)) & 1 else c[dd] for dd in range(3)]
            nb[sub] = (cc[0] * tiles[1] + cc[1]) * tiles[2] + cc[2]
            assert 0 <= nb[sub] < ntiles
        for q in range(PT**3):
            i, j, k = q >> (2 * PTB), (q >> PTB) & (PT - 1), q & (PT - 1)
            if i >= e[0] or j >= e[1] or k >= e[2]: continue
            zero = (4 if i == 0 else 0) | (2 if j == 0 else 0) | (1 if k == 0 else 0)
            if zero == 0: continue
            total, any_ = 0., False
            if full[nb[0]]:
                v = bnd[nb[0], (i * PTH + j) * PTH + k]
                assert np.isfinite(v), 'own boundary cell never written'
                total, any_ = v, True
            for sub in range(1, 8):
                if (sub & zero) != sub or not full[nb[sub]]: continue
                ii = em[0] if sub & 4 else i
                jj = em[1] if sub & 2 else j
                kk = em[2] if sub & 1 else k
                v = bnd[nb[sub], (ii * PTH + jj) * PTH + kk]
                assert np.isfinite(v), 'neighbour halo cell never written'
                total += v
                any_ = True
            g = (o[0] + i, o[1] + j, o[2] + k)
            if any_: mesh[g] += total
            writes[g] += 1
    return mesh, writes


def direct_cic(shape, p, w):
    mesh = np.zeros(shape)
    b = np.floor(p).astype(int) % np.array(shape)
    d = p - b
    for a, bb, c in itertools.product((0, 1), repeat=3):
        wt = (d[:, 0] if a else 1 - d[:, 0]) * (d[:, 1] if bb else 1 - d[:, 1]) * (d[:, 2] if c else 1 - d[:, 2])
        idx = tuple((b[:, k] + s) % n for k, (s, n) in enumerate(zip((a, bb, c), shape)))
        np.add.at(mesh, idx, wt * w)
    return mesh


@pytest.mark.parametrize('shape, extent', [((16, 16, 16), 1.), ((12, 20, 9), 1.), ((4, 8, 3), 1.), ((1, 2, 17), 1.),
                                           ((8, 8, 8), 1.), ((24, 5, 16), 1.), ((32, 24, 16), 0.3), ((40, 34, 18), 1.),
                                           ((17, 33, 48), 0.5)])
def test_tile_decomposition(shape, extent):
    rng = np.random.RandomState(sum(shape))
    n = 400
    p = rng.uniform(0., extent, (n, 3)) * np.array(shape)  # extent < 1: empty tiles
    # points on cell and tile faces and at the last cell (periodic wrap of the +1 corner)
    p[:10] = np.floor(p[:10])
    p[10:20, 0] = shape[0] - 1e-3
    w = rng.uniform(0.5, 1.5, n)
    mesh, writes = model_paint(shape, p, w)
    assert writes.max() <= 1, 'a cell is written by both passes'
    assert np.allclose(mesh, direct_cic(shape, p, w), rtol=1e-12, atol=1e-12)
