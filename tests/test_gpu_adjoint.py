"""The discrete adjoint (reverse mode, checkpointed) of the step loop against the forward-mode
derivative of the same discrete map (nf_enable_sensitivity): both are exact derivatives of the
same floating-point trajectory, so their twelve gradients of (L/L0, mushroom diameter) with
respect to (friction, A, B, n, C, m) agree to rounding. Also: the forward trajectory with the
adjoint recording is the plain one bit for bit, the sweep leaves the solver in its final state,
and the checkpoint interval changes no bit of the gradient."""
import numpy as np
import pytest

from conftest import same_bits
from neofiz_amd import sensitivity as S, taylor
from neofiz_amd.solver import Contact, IllegalStateException

pytestmark = pytest.mark.gpu

JC = taylor.ROD_JC.johnsonCook


def point(nr, nz, speed=taylor.ROD_SPEED, mu=taylor.ROD_JC_FRICTION):
    return S.Point(nr, nz, speed, mu, JC.a, JC.b, JC.n, JC.c, JC.m)


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


@pytest.mark.parametrize("nr,nz,end,tol", [(16, 160, 5e-6, 1e-10), (16, 160, 20e-6, 1e-10),
                                           (16, 160, 80e-6, 1e-9), (32, 320, 80e-6, 1e-9)])
def test_adjoint_equals_the_tangent(nr, nz, end, tol):
    pt = point(nr, nz)
    print(f"[adjoint] {nr}x{nz} to {end * 1e6:.0f} us", flush=True)
    obj_t, grad_t, steps_t, _ = S.gradients(pt, end)
    obj_a, grad_a, steps_a, fwd, rev, info = S.adjoint_gradients(pt, end)
    assert steps_a == steps_t and obj_a == obj_t
    err = rel(grad_a, grad_t)
    print(f"[adjoint] steps {steps_a}, checkpoints {info[1]}, capped solves {info[2]}, "
          f"max rel {err.max():.2e}, forward {fwd:.2f} s, reverse {rev:.2f} s", flush=True)
    assert np.all(grad_t != 0.0)
    assert err.max() <= tol, (grad_a, grad_t)


def test_recording_keeps_the_trajectory_and_the_sweep_restores_it():
    pt = point(8, 80)
    mesh, plain = S.build(pt)
    plain.runUntil(10e-6)
    _, rec = S.build(pt)
    rec.enableAdjoint(16)
    rec.runUntil(10e-6)
    fields = ("ur", "uz", "vr", "vz", "fr", "fz")
    for k in fields:
        assert same_bits(rec.nodeField(k), plain.nodeField(k)), k
    assert same_bits(rec.stress(), plain.stress()) and rec.steps() == plain.steps()
    _, ext = S.objectives(mesh, rec)
    rec.adjointGradient(S.objective_seeds(mesh, ext))
    for k in fields:
        assert same_bits(rec.nodeField(k), plain.nodeField(k)), k
    assert same_bits(rec.plasticStrain(), plain.plasticStrain()) and rec.time() == plain.time()
    # and the run continues from there exactly as the plain one does
    rec.run(37)
    plain.run(37)
    assert same_bits(rec.axialDisplacement(), plain.axialDisplacement())
    plain.shutdown()
    rec.shutdown()


def test_checkpoint_interval_changes_no_bit():
    pt = point(8, 80)
    grads = [S.adjoint_gradients(pt, 20e-6, interval=k)[1] for k in (64, 5, 1000)]
    assert same_bits(grads[0], grads[1]) and same_bits(grads[0], grads[2])


def test_adjoint_refusals():
    pt = point(4, 40)
    mesh, s = S.build(pt)
    s.run(3)
    with pytest.raises(IllegalStateException):
        s.enableAdjoint(64)
    s.shutdown()
    mesh, s = S.build(pt)
    s.enableAdjoint(64)
    with pytest.raises(IllegalStateException):
        s.setContact(Contact(mesh))
    with pytest.raises(IllegalStateException):
        s.enableSensitivity()
    with pytest.raises(IllegalStateException):
        s.adjointGradient(np.zeros((1, 4, mesh.nodeCount)))  # nothing recorded yet
    s.shutdown()
