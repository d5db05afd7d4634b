"""The fast f32 kernels on 2x2x2-element blocks against the two-element-tile
kernels: the same bits after several steps, for every scheme, Euler and
Navier-Stokes, walls and periodic y.

A box with an odd dimension stepped alone runs on tiles; in a one-box
ensemble its core runs on blocks and only its last layers on tiles, so the
two runs compare the block and the tile kernels element by element.  The
thinnest boxes (2 elements along two directions) give a block that is its
own periodic neighbour across those faces.  A box with all dimensions even
runs the block kernels in both forms (the single-box launch and the
ensemble's), which must agree as well."""
import numpy as np
import pytest

import opendynamo_amd as od
from helpers import smooth_state

pytestmark = pytest.mark.gpu

ODD_DIMS = [(3, 2, 2, 3.0, 2.0, 2.0), (2, 3, 2, 2.0, 3.0, 2.0), (2, 2, 3, 2.0, 2.0, 3.0),
            (5, 6, 3, 5.0, 2.0, 3.0), (7, 4, 9, 6.0, 2.0, 4.0)]
EVEN_DIMS = [(2, 2, 2, 6.283185307179586, 2.0, 3.141592653589793), (4, 6, 2, 6.0, 2.0, 3.0),
             (6, 4, 8, 4.0, 2.0, 5.0)]
DIMS = ODD_DIMS + EVEN_DIMS
FLUX = ["slau2", "ausm+", "rusanov"]


def announce(**case):
    print("CASE", " ".join(f"{k}={v}" for k, v in case.items()), flush=True)


@pytest.mark.parametrize("dims", DIMS, ids=lambda d: f"{d[0]}x{d[1]}x{d[2]}")
@pytest.mark.parametrize("flux", FLUX)
def test_block_kernels_equal_tile_kernels(dims, flux):
    nx, ny, nz, lx, ly, lz = dims
    for viscous in (True, False):
        for wall in (True, False):
            announce(test="block_vs_tile", dims=dims[:3], flux=flux, viscous=viscous, wall=wall)
            kw = dict(u_real=od.OD_F32, flux=flux, y_wall=wall, viscous=viscous, viscosity_model=1,
                      mu_ref=0.002 if viscous else 0.0, fast_math=1)
            u0 = smooth_state(nx, ny, nz, lx, ly, lz, seed=nx + 3 * ny + 7 * nz, amp=0.05, dtype=np.float32)
            dt = 0.002
            with od.Solver(od.params(nx, ny, nz, lx, ly, lz, **kw)) as s:
                s.set_state(u0)
                s.step(dt, 4)
                got = s.get_state()
            with od.Batch(od.params(1, 1, 1, 1.0, 1.0, 1.0, **kw), [dims]) as b:
                b.set_schedule(orient=0, pairs=1, order=0, groups_per_cu=0)
                b.set_state(0, u0)
                b.step([dt], 4)
                ref = b.get_state(0)
            assert np.isfinite(got).all()
            assert got.tobytes() == ref.tobytes(), (dims, flux, viscous, wall)
