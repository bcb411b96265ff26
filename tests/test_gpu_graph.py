"""HIP-graph replays of OS-SQS passes and of the learned unrolled
reconstruction equal the eager launches bit for bit, on every replay, with
unrelated eager GPU work and synchronisations between the replays."""

import numpy as np
import pytest
import torch

from reconnet_amd.phantom import phantom
from tests.helpers import small_fan

pytestmark = pytest.mark.gpu


def _engine(nslices=2, nx=96, nview=96, nsub=8):
    from reconnet_amd.sqs import OsSqsTV
    from reconnet_amd.torch_ops import FanPlan

    geo = small_fan(nx=nx, nu=160, nview=nview, dx=1.5)
    imgs = torch.stack([torch.from_numpy(phantom(s, nx, nx, geo.dx, geo.dy)) for s in range(nslices)]).cuda()
    sino = FanPlan(geo).forward(imgs, "siddon")
    eng = OsSqsTV(geo, nsub, nslices)
    eng.set_data(sino, torch.ones_like(sino))
    eng.compute_norm()
    return eng


def _noise_work():
    a = torch.randn(1 << 20, device="cuda")
    (a * 2 + 1).sum().item()
    torch.cuda.synchronize()


@pytest.mark.parametrize("nslices", [2, 9])
def test_pass_graph_replays_equal_eager(nslices):
    # 9 slices: the engine's interleaved layout (three 4-slice groups)
    eng = _engine(nslices)
    x0 = torch.full((eng.nslices, eng.geo.nx, eng.geo.ny), 0.01, device="cuda")
    eng.set_image(x0)
    for _ in range(3):
        eng.run_pass(beta=1e-4, eps=1e-6, lb=0.0)
    ref = eng.image().clone()
    eng.set_image(x0)
    pg = eng.capture_passes(3, beta=1e-4, eps=1e-6, lb=0.0)
    assert torch.equal(eng.image(), x0)
    for _ in range(3):
        eng.set_image(x0)
        pg.replay()
        torch.cuda.synchronize()
        assert torch.equal(eng.image(), ref)
        _noise_work()


def test_pass_graph_rejects_odd_step_count():
    eng = _engine(nsub=3, nview=96)
    with pytest.raises(ValueError):
        eng.capture_passes(1)


@pytest.mark.parametrize("nslices", [2, 9])
def test_unrolled_graph_replays_equal_eager(nslices):
    from reconnet_amd.cnn import PriorCNN, init_params
    from reconnet_amd.unrolled import UnrolledGraph, UnrolledOSSQS

    eng = _engine(nslices)
    nets = [PriorCNN(init_params(seed=k, last_scale=0.5), scale=50.0) for k in range(2)]
    model = UnrolledOSSQS(nets)
    ref = model.reconstruct(eng).clone()
    ug = UnrolledGraph(model, eng)
    for _ in range(3):
        out = ug.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        _noise_work()
        # the engine's device buffers are shared with the graph: eager use in between is fine
        model.reconstruct(eng)
        torch.cuda.synchronize()
