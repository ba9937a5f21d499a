"""The ring Jacobi kernel's tile layout (RingLayout, firex_jacobi.hip):
ring_tiles, ring_tile and ring_layout restated here.  With the 32-column
wall tile, every output column of a row is written by exactly one tile, and
each tile holds its output columns at least K cells away from any tile edge
that is not a wall (validity shrinks by one cell per level from such edges).
The GPU parity tests run the device code on grids with and without the wall
tile; this checks the layout arithmetic for every width the kernel takes."""
import pytest


def ring_tiles(n, T, K, top):
    span = n + K + (0 if top else K)
    return 1 if span <= T else 1 + (span - T + (T - 2 * K) - 1) // (T - 2 * K)


def ring_tile(i, count, n, T, K, top):
    if i == count - 1:
        origin, hi = n + (0 if top else K) - T, n
    else:
        origin = -K + i * (T - 2 * K)
        hi = origin + T - K
    return origin, (0 if i == 0 else origin + K), hi


def ring_layout(nx, ny, K, H, lane_rows):
    tiles_y = ring_tiles(ny, H, K, 1)
    step = 64 - 2 * K
    full, left = nx // step, nx - (nx // step) * step
    if lane_rows and full >= 1 and 0 < left <= 32 - 2 * K:
        return full, tiles_y, 1, (tiles_y + 1) // 2
    return ring_tiles(nx, 64, K, 0), tiles_y, 0, 0


def x_tiles(nx, K, lane_rows=True):
    """(origin, width, lo, hi) of each x tile."""
    tx, _, narrow, _ = ring_layout(nx, 64, K, 32, lane_rows)
    out = [(o, 64, lo, hi) for o, lo, hi in (ring_tile(i, tx + narrow, nx, 64, K, 0) for i in range(tx))]
    if narrow:
        out.append((nx + K - 32, 32, tx * (64 - 2 * K), nx))
    return out


@pytest.mark.parametrize("K", [1, 2, 3, 4, 5])
def test_wall_tile_covers_each_column_once_with_valid_cells(K):
    for nx in list(range(64, 300)) + [508, 1024]:
        tiles = x_tiles(nx, K)
        cover = [0] * nx
        for o, w, lo, hi in tiles:
            assert o >= -K and o + w <= nx + K, (nx, K, o, w)  # mirror halo reaches K cells past each wall
            for x in range(lo, hi):
                cover[x] += 1
                left_ok = o == -K or x - o >= K
                right_ok = o + w == nx + K or (o + w - 1) - x >= K
                assert left_ok and right_ok, (nx, K, o, w, x)
        if ring_layout(nx, 64, K, 32, True)[2]:
            assert cover == [1] * nx, (nx, K)
        else:
            assert min(cover) >= 1, (nx, K)


def test_wall_tile_counts_at_the_benchmark_widths():
    # 192 columns at K = 4: three 64-column tiles and the wall tile (224
    # columns computed per row instead of 256); 42 tile columns x 12 chunks
    # of 16 planes = 504 blocks
    assert ring_layout(192, 288, 4, 32, True) == (3, 12, 1, 6)
    assert sum(w for _, w, _, _ in x_tiles(192, 4)) == 224
    assert sum(w for _, w, _, _ in x_tiles(192, 4, lane_rows=False)) == 256
    tx, ty, narrow, pairs = ring_layout(192, 288, 4, 32, True)
    assert (tx * ty + pairs) * ((192 + 15) // 16) == 504
    assert ring_layout(128, 192, 4, 32, True)[2] == 1     # HD: 160 instead of 192 columns
    assert ring_layout(160, 240, 4, 32, True)[2] == 0     # 48 columns left: a full tile
    assert ring_layout(192, 288, 4, 32, False)[2] == 0    # 8-wave blocks: no wall tile
