"""The hand-written adjoint against torch's reverse mode through an FP64
brute-force restatement (oracle/torch_ref.py) and against central finite
differences (h = 1e-6) of the engine's own forward loss."""

import numpy as np
import pytest
import torch

from oracle import ljref
from oracle import torch_ref

pytestmark = pytest.mark.gpu


def _engine():
    from lennardjonesium import engine as eng
    assert eng.device_count() > 0
    return eng


def _loss(eng, eps, sig, pos, vel, L, nsteps, rho, cutoff=2.5, dt=0.005):
    with eng.Engine([eng.SystemSpec(len(pos), L, cutoff=cutoff, time_delta=dt, epsilon=eps,
                                    sigma=sig)]) as e:
        e.set_state(0, pos, vel)
        return e.adjoint(nsteps, rho, every=7)


@pytest.mark.parametrize("eps,sig", [(1.0, 1.0), (1.3, 0.95)])
def test_adjoint_matches_torch_reverse_mode(eps, sig):
    eng = _engine()
    st, L = ljref.initial_condition(108, 0.8, 1.0, seed=7)
    L = L * sig
    pos, vel = st.pos * sig, st.vel
    nsteps, rho = 20, 0.8
    loss, ge, gs, gr, gv = _loss(eng, eps, sig, pos, vel, L, nsteps, rho)
    te = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    ts = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    tr = torch.tensor(pos, requires_grad=True)
    tv = torch.tensor(vel, requires_grad=True)
    ref = torch_ref.trajectory_loss(te, ts, tr, tv, L, nsteps, 0.005, 2.5, rho)
    ref.backward()
    print("loss", loss, ref.item(), "eps", ge, te.grad.item(), "sig", gs, ts.grad.item(), flush=True)
    assert abs(loss - ref.item()) < 1e-11 * abs(ref.item())
    assert abs(ge - te.grad.item()) < 1e-8 * abs(te.grad.item())
    assert abs(gs - ts.grad.item()) < 1e-8 * abs(ts.grad.item())
    scale = np.abs(tr.grad.numpy()).max()
    assert np.abs(gr - tr.grad.numpy()).max() < 1e-8 * scale
    assert np.abs(gv - tv.grad.numpy()).max() < 1e-8 * np.abs(tv.grad.numpy()).max()


def test_adjoint_matches_central_differences():
    eng = _engine()
    pos, vel, L = eng.initial_condition(500, 0.8, 1.0, seed=3)
    rho, nsteps, h = 0.8, 100, 1e-6
    loss, ge, gs, gr, gv = _loss(eng, 1.0, 1.0, pos, vel, L, nsteps, rho)

    def fwd(eps=1.0, sig=1.0, p=pos, v=vel, box=L):
        return _loss(eng, eps, sig, p, v, box, nsteps, rho)[0]

    fd_e = (fwd(eps=1 + h) - fwd(eps=1 - h)) / (2 * h)
    # sigma scales the pair law and the cutoff, not the box or the positions
    fd_s = (fwd(sig=1 + h) - fwd(sig=1 - h)) / (2 * h)
    print("eps", ge, fd_e, "sig", gs, fd_s, flush=True)
    assert abs(ge - fd_e) < 1e-4 * abs(fd_e)
    assert abs(gs - fd_s) < 1e-4 * abs(fd_s)
    rng = np.random.default_rng(0)
    for i, a in zip(rng.integers(0, 500, 4), rng.integers(0, 3, 4)):
        pp, pm = pos.copy(), pos.copy()
        pp[i, a] += h
        pm[i, a] -= h
        fd = (fwd(p=pp) - fwd(p=pm)) / (2 * h)
        assert abs(gr[i, a] - fd) < 1e-4 * max(abs(fd), 1e-3 * np.abs(gr).max())


def test_torch_autograd_function_wraps_the_adjoint():
    eng = _engine()
    from lennardjonesium.autograd import trajectory_loss
    pos, vel, L = eng.initial_condition(256, 0.8, 1.0, seed=5)
    eps = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    r0 = torch.tensor(pos, requires_grad=True)
    v0 = torch.tensor(vel, requires_grad=True)
    loss = trajectory_loss(eps, sig, r0, v0, L, nsteps=50, density=0.8, every=10)
    (2.0 * loss).backward()
    l2, ge, gs, gr, gv = _loss(eng, 1.0, 1.0, pos, vel, L, 50, 0.8)
    assert loss.item() == l2
    assert eps.grad.item() == pytest.approx(2 * ge, rel=1e-12)
    assert sig.grad.item() == pytest.approx(2 * gs, rel=1e-12)
    np.testing.assert_allclose(r0.grad.numpy(), 2 * gr, rtol=1e-12, atol=1e-14)


def test_adjoint_mid_run_keeps_displacements_and_time():
    """Called after steps, the adjoint leaves the engine with the state it
    found -- positions, velocities, displacements, elapsed time -- and its
    gradients equal those of a fresh engine loaded with that state (the loss
    does not depend on displacements or time)."""
    eng = _engine()
    pos, vel, L = eng.initial_condition(500, 0.8, 1.0, seed=4)
    with eng.Engine([eng.SystemSpec(500, L)]) as e:
        e.set_state(0, pos, vel)
        e.step(13)
        before = e.get_state(0)
        assert before["time"] > 0 and np.abs(before["disp"]).max() > 0
        got = e.adjoint(30, 0.8, every=10)
        after = e.get_state(0)
        rec = e.step(5)
    for key in ("pos", "vel", "disp"):
        assert np.array_equal(after[key], before[key]), key
    assert after["time"] == before["time"]
    with eng.Engine([eng.SystemSpec(500, L)]) as f:
        f.set_state(0, before["pos"], before["vel"])
        want = f.adjoint(30, 0.8, every=10)
    assert got[0] == want[0] and got[1] == want[1] and got[2] == want[2]
    assert np.array_equal(got[3], want[3])
    # the run continues from the kept clock and displacements
    st = ljref.State(pos=before["pos"], vel=before["vel"])
    st.disp = before["disp"].copy()
    st.time = before["time"]
    st.F, st.pe, st.vir = ljref.forces(before["pos"], L)
    ref = ljref.run(st, L, 0.005, 5)
    assert np.abs(rec[:, 0, :] - ref).max() <= 1e-10 * np.abs(ref).max()
