"""torch.autograd binding of the adjoint form (cloverleaf_amd.autograd) on the GPU:
the gradients torch receives are exactly the C ABI's (clv_adjoint_backward),
for CUDA and CPU tensors, scaled by the incoming gradient, and agree with
central differences of the forward pass."""
import numpy as np
import pytest

import cloverleaf_amd as clv

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _setup(nx, ny):
    ch = clv.Chunk(nx, ny)
    ch.start(clv.parse_deck(clv.bm_deck(nx, ny)))
    rho0 = ch.download("density0")[2:ny + 2, 2:nx + 2]
    e0 = ch.download("energy0")[2:ny + 2, 2:nx + 2]
    rho0, e0 = clv.adjoint_initial(rho0, e0, 1e-2, 3)
    return ch, rho0, e0 * (1 + 1e-2 * clv.smooth_field(nx, ny, 5))


def test_autograd_matches_the_adjoint_and_differences(gpu):
    from cloverleaf_amd.autograd import kinetic_energy
    nx, ny, steps, dt = 48, 40, 10, 0.01
    ch, rho0, e0 = _setup(nx, ny)
    with clv.Adjoint(ch, dt, steps, 4) as adj:
        rho = torch.tensor(rho0, device="cuda", requires_grad=True)
        e = torch.tensor(e0, device="cuda", requires_grad=True)
        ke = kinetic_energy(adj, rho, e)
        (2.0 * ke).backward()
        adj.set_initial(rho0, e0)
        s = adj.forward()
        assert float(ke.detach()) == s[3]
        gr, ge = adj.backward(2.0)
        assert np.array_equal(rho.grad.cpu().numpy(), gr) and np.array_equal(e.grad.cpu().numpy(), ge)
        # host tensors
        rc = torch.tensor(rho0, requires_grad=True)
        ec = torch.tensor(e0, requires_grad=True)
        kc = kinetic_energy(adj, rc, ec)
        kc.backward()
        assert float(kc.detach()) == s[3]
        assert np.array_equal(rc.grad.numpy() * 2.0, gr) and np.array_equal(ec.grad.numpy() * 2.0, ge)
        # central differences along a smooth direction
        d = clv.smooth_field(nx, ny, 11)
        eps = 1e-7
        with torch.no_grad():
            kp = float(kinetic_energy(adj, torch.tensor(rho0 + eps * d, device="cuda"), torch.tensor(e0, device="cuda")))
            km = float(kinetic_energy(adj, torch.tensor(rho0 - eps * d, device="cuda"), torch.tensor(e0, device="cuda")))
        fd = (kp - km) / (2 * eps)
        an = float(np.sum(rc.grad.numpy() * d))
        assert abs(an - fd) <= 1e-6 * abs(fd), (an, fd)


def test_autograd_backward_after_another_forward(gpu):
    """Two forwards on one Adjoint, then backward through the first: the gradient
    is the first input's (its forward is re-run from the saved inputs)."""
    from cloverleaf_amd.autograd import kinetic_energy
    nx, ny, steps, dt = 32, 24, 6, 0.01
    ch, rho0, e0 = _setup(nx, ny)
    with clv.Adjoint(ch, dt, steps, 3) as adj:
        adj.set_initial(rho0, e0)
        adj.forward()
        gr, ge = adj.backward(1.0)
        rho = torch.tensor(rho0, device="cuda", requires_grad=True)
        e = torch.tensor(e0, device="cuda", requires_grad=True)
        ke1 = kinetic_energy(adj, rho, e)
        rho2 = torch.tensor(rho0 * 1.01, device="cuda", requires_grad=True)
        ke2 = kinetic_energy(adj, rho2, torch.tensor(e0, device="cuda"))
        assert float(ke2.detach()) != float(ke1.detach())
        ke1.backward(retain_graph=True)
        assert np.array_equal(rho.grad.cpu().numpy(), gr) and np.array_equal(e.grad.cpu().numpy(), ge)
        # a second backward through the same graph: the first consumed the forward run
        rho.grad = None
        e.grad = None
        ke1.backward()
        assert np.array_equal(rho.grad.cpu().numpy(), gr) and np.array_equal(e.grad.cpu().numpy(), ge)


def test_autograd_rejects_wrong_shapes_and_types(gpu):
    from cloverleaf_amd.autograd import kinetic_energy
    ch, rho0, e0 = _setup(16, 12)
    with clv.Adjoint(ch, 0.01, 2, 1) as adj:
        with pytest.raises(ValueError):
            kinetic_energy(adj, torch.zeros((12, 15), dtype=torch.float64), torch.zeros((12, 16), dtype=torch.float64))
        with pytest.raises(TypeError):
            kinetic_energy(adj, torch.zeros((12, 16)), torch.zeros((12, 16)))
