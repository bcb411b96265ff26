"""The projectors as torch.library operators: schema / fake-tensor / autograd
registration checks (torch.library.opcheck), exact-adjoint gradients, and a
graph compiled with the AOT-autograd eager backend that matches eager mode."""

import numpy as np
import pytest
import torch

from tests.helpers import small_fan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fan():
    from reconnet_amd import ops
    from reconnet_amd.torch_ops import FanPlan

    geo = small_fan(nx=48, nu=80, nview=60)
    pid = ops.register_plan(FanPlan(geo))
    yield geo, pid
    ops.release_plan(pid)


@pytest.mark.parametrize("kind", ["dd", "siddon"])
def test_fan_ops_opcheck(fan, kind):
    geo, pid = fan
    x = torch.rand(2, geo.nx, geo.ny, device="cuda", requires_grad=True)
    torch.library.opcheck(torch.ops.reconnet.fan_forward.default, (x, pid, kind))
    y = torch.rand(2, geo.nu, geo.nview, device="cuda", requires_grad=True)
    adj = {"dd": "dd_adjoint", "siddon": "siddon"}[kind]
    torch.library.opcheck(torch.ops.reconnet.fan_backward.default, (y, pid, adj))


def test_fan_op_gradient_is_exact_adjoint(fan):
    from reconnet_amd import ops

    geo, pid = fan
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(2, geo.nx, geo.ny, device="cuda", generator=g, requires_grad=True)
    v = torch.rand(2, geo.nu, geo.nview, device="cuda", generator=g)
    y = ops.fan_project(x, pid, "dd")
    (y * v).sum().backward()
    lhs = float((y.detach().double() * v.double()).sum())
    rhs = float((x.detach().double() * x.grad.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5


def test_compiled_graph_matches_eager(fan):
    from reconnet_amd import ops

    geo, pid = fan
    meas = torch.rand(1, geo.nu, geo.nview, device="cuda")

    def loss(x):
        r = ops.fan_project(torch.relu(x) * 2.0, pid, "siddon") - meas
        return 0.5 * (r * r).sum()

    x0 = torch.rand(1, geo.nx, geo.ny, device="cuda")
    xe = x0.clone().requires_grad_(True)
    le = loss(xe)
    le.backward()
    xc = x0.clone().requires_grad_(True)
    lc = torch.compile(loss, backend="aot_eager", fullgraph=True)(xc)
    lc.backward()
    assert torch.allclose(lc, le, rtol=1e-6)
    assert torch.equal(xc.grad, xe.grad)


def test_cone_ops_opcheck_and_adjoint():
    from reconnet_amd import ops
    from reconnet_amd.torch_ops import ConePlan
    from tests.test_gpu_cone import circ_cone

    plan = ConePlan(16, 16, 8, 1.0, 1.0, 1.0, 24, 12, 1.5, 1.5, *circ_cone(12, 150.0, 250.0))
    pid = ops.register_plan(plan)
    try:
        x = torch.rand(1, 16, 16, 8, 1, device="cuda", requires_grad=True)
        torch.library.opcheck(torch.ops.reconnet.cone_forward.default, (x, pid, "dd"))
        y = ops.cone_project(x, pid, "dd")
        v = torch.rand_like(y)
        (y * v).sum().backward()
        lhs = float((y.detach().double() * v.double()).sum())
        rhs = float((x.detach().double() * x.grad.double()).sum())
        assert abs(lhs - rhs) / abs(lhs) <= 1e-5
    finally:
        ops.release_plan(pid)


@pytest.mark.parametrize("kind", ["dd", "siddon"])
def test_fan_ops_values_and_gradient_match_oracle(fan, kind):
    # the TF op contract (zOpsReconNet.cc:270-321, ReconNet.py:773-830): the
    # forward op's values, and the gradient of 0.5 |A x - y|^2 taken through the
    # op's registered backward, A^T (A x - y), against the FP64 oracle (for DD
    # the oracle's A^T is the explicit transpose of its DD FP on this grid)
    import oracle as orc
    from reconnet_amd import ops
    from tests.helpers import rel_l2

    geo, pid = fan
    deg = geo.angles()
    g5 = geo.oracle_geo()
    rng = np.random.default_rng(17)
    x0 = rng.uniform(0, 1, (2, geo.nx, geo.ny)).astype(np.float32)
    meas = rng.uniform(0, 20, (2, geo.nu, geo.nview)).astype(np.float32)
    x = torch.from_numpy(x0).cuda().requires_grad_(True)
    y = ops.fan_project(x, pid, kind)
    r = y - torch.from_numpy(meas).cuda()
    (0.5 * (r * r).sum()).backward()
    fp = orc.dd_fan_fp if kind == "dd" else orc.siddon_fan_fp
    ref_y = fp(x0[..., None, None], deg, g5, geo.nu)[..., 0, 0]
    assert rel_l2(y.detach().cpu().numpy(), ref_y) <= 1e-5
    res = (ref_y - meas).astype(np.float64)
    if kind == "siddon":
        ref_g = orc.siddon_fan_bp(res.astype(np.float32)[..., None, None], deg, g5, (geo.nx, geo.ny, 1))[..., 0, 0]
        # the residual fed to the oracle is rounded to float32 once
        tol = 1e-5
    else:
        cols = []
        for k in range(geo.nx * geo.ny):
            e = np.zeros(geo.nx * geo.ny, np.float32)
            e[k] = 1
            cols.append(orc.dd_fan_fp(e.reshape(1, geo.nx, geo.ny, 1, 1), deg, g5, geo.nu).ravel())
        A = np.stack(cols, 1)
        ref_g = np.stack([(A.T @ res[b].ravel()).reshape(geo.nx, geo.ny) for b in range(2)])
        tol = 1e-5
    assert rel_l2(x.grad.cpu().numpy(), ref_g) <= tol
