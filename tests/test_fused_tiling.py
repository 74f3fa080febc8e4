"""Geometry of the fused one-launch step (fused_step_kernel in csrc/swe_stage.hip), checked on
the CPU: with 52-column strips (lane l on column c0 - 6 + l) and row segments of T rows, a
tile's stage 1 writes lanes [2, 62) over rows [r0-4, r1+4), stage 2 lanes [4, 60) over
[r0-2, r1+2), stage 3 lanes [6, 58) over [r0, r1) (interior cells only).  Every cell that
an owner of stage s+1 reads through the radius-2 cross stencil (ghost rows and columns
mirrored to the interior, for both wall types) must have been written by the same tile in
stage s, and the stage-3 outputs must tile the interior exactly once."""
import itertools

import pytest

ng = 2


def mirror(r, lo, hi, bc):
    """Source cell of row/column r (up to 2 outside [lo, hi)): reflecting or transmissive."""
    if r < lo:
        return 2 * lo - 1 - r if bc == 0 else lo
    if r >= hi:
        return 2 * hi - 1 - r if bc == 0 else hi - 1
    return r


def tiles(nx, ny, T, bc):
    """Yield each tile's stage-3 output cells after checking its stage dependencies."""
    strips = (ny + 51) // 52
    segs = (nx + T - 1) // T
    ie, je = ng + nx, ng + ny
    for strip, seg in itertools.product(range(strips), range(segs)):
        col0 = ng - 6 + strip * 52
        r0 = ng + seg * T
        r1 = min(r0 + T, ie)

        def owned(lo, hi, ra, rb):
            return {(r, col0 + l) for l in range(lo, hi) for r in range(ra, rb)
                    if ng <= col0 + l < je}
        w1 = owned(2, 62, max(r0 - 4, ng), min(r1 + 4, ie))
        w2 = owned(4, 60, max(r0 - 2, ng), min(r1 + 2, ie))
        w3 = owned(6, 58, r0, r1)
        for (wprev, wcur) in ((w1, w2), (w2, w3)):
            for (r, j) in wcur:
                for dr in range(-2, 3):
                    for dj in range(-2, 3):
                        if dr and dj:
                            continue  # cross stencil
                        rr = mirror(r + dr, ng, ie, bc)
                        jj = mirror(j + dj, ng, je, bc)
                        assert (rr, jj) in wprev, (nx, ny, T, strip, seg, r, j, dr, dj)
        # outputs tile the interior exactly once
        yield w3
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("T", [4, 5, 8, 64])
def test_fused_tiles_read_only_their_own_stage_output(T, bc):
    for nx, ny in itertools.product([2, 3, 5, 17, 53, 60, 105], [2, 3, 5, 51, 52, 53, 104, 105, 130]):
        cells = set()
        for w3 in tiles(nx, ny, T, bc):
            assert not (cells & w3)
            cells |= w3
        assert cells == {(r, j) for r in range(ng, ng + nx) for j in range(ng, ng + ny)}, (nx, ny, T)
