"""GPU parity of the GP noise: cell impulses bit-exact, psi / grad psi within 1e-5 of FP64."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import kernels, random_points

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ki", range(len(kernels())))
def test_cell_impulses_bit_exact(ctx, dev, ki):
    import torch
    from mupsi_amd import batch
    k = kernels()[ki]
    rng = np.random.default_rng(100 + ki)
    n = 4096
    cells = rng.integers(-3000, 3000, (n, 3)).astype(np.int32)
    cells[:8] = [[0, 0, 0], [-1, -1, -1], [2**31 - 1, -2**31, 0], [-2**31, 5, 2**31 - 1],
                 [1, 0, -1], [100, -100, 7], [0, 1, 0], [-1, 0, 1]]
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    cs_o, pos_o, w_o = O.cell_impulses(k, cells, seeds)
    cs, pos, w = batch.cell_impulses(ctx, k, torch.from_numpy(cells).to(dev),
                                     torch.from_numpy(seeds.view(np.int32)).to(dev))
    assert np.array_equal(cs.cpu().numpy().view(np.uint32), cs_o)
    assert np.array_equal(pos.cpu().numpy().view(np.uint32), pos_o.view(np.uint32))
    assert np.array_equal(w.cpu().numpy(), w_o)


@pytest.mark.parametrize("ki", range(len(kernels())))
def test_psi_within_1e5_of_fp64(ctx, dev, ki):
    import torch
    from mupsi_amd import batch
    k = kernels()[ki]
    rng = np.random.default_rng(200 + ki)
    n = 20000
    p = random_points(rng, n, 600.0 * k["length_scale"])
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    ref, l1 = O.psi_eval(k, p, seeds, fp64=True)
    out, l1_gpu = batch.psi_eval(ctx, k, torch.from_numpy(p).to(dev),
                                 torch.from_numpy(seeds.view(np.int32)).to(dev))
    got = out.cpu().numpy().astype(np.float64)
    # L1-mass-normalized tolerance: |a - b| <= 1e-5 * sigma/sqrt(Var) * sum |h_i|
    L2 = float(np.sum(np.square(k["aniso"])))
    gscale = l1 * (1.0 + 2.0 / L2 * 3.0 * k["length_scale"] * 1.8)  # bounds |grad h| mass
    err_v = np.abs(got[:, 0] - ref[:, 0])
    err_g = np.abs(got[:, 1:] - ref[:, 1:]).max(axis=1)
    assert np.all(err_v <= 1e-5 * l1 + 1e-30), err_v.max()
    assert np.all(err_g <= 1e-5 * gscale + 1e-30), err_g.max()
    np.testing.assert_allclose(l1_gpu.cpu().numpy(), l1, rtol=1e-5)
