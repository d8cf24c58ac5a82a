"""torch.autograd bridge to the discrete adjoint of the LEE time advance.

The functional is the observer energy of the C ABI (lee_set_observer):
J = sum over steps k in (window_start, nsteps] of sum_l w_l p'_k(l)^2.
Its inputs are the source field q of a SOURCE_FIELD block (p-equation
source q(x) A sin(omega t)) and the initial state y0.  The forward pass
runs lee_objective_forward (state checkpointed every `ckpt_every` steps on
the device); the backward pass runs lee_objective_backward, which
recomputes each checkpoint segment and runs the transposed step kernels
through it, and returns dJ/dq and dJ/dy0.

torch carries the tensors only: the forward and reverse time advance run
in the HIP library.
"""
import torch

from .lee import Solver  # noqa: F401  (type of the solver argument)


def _run_forward(solver, q, y0, clock, nsteps, window_start, ckpt_every):
    solver.set_source_field(q)
    solver.set_solution(y0)
    solver.set_time(*clock)
    return solver.objective_forward(nsteps, window_start, ckpt_every)


class ObserverEnergy(torch.autograd.Function):
    """The solver keeps the checkpoints of its latest forward run only.  The
    context therefore records the run's inputs, observers and the solver's
    objective generation; when another forward run (or a change of
    observers) has happened on the same solver since, backward first re-runs
    this forward from the recorded inputs, so J1 + J2 on one solver, or a
    finite-difference probe between forward and backward, still get the
    gradient of their own run."""

    @staticmethod
    def forward(ctx, q, y0, solver, nsteps, window_start, ckpt_every, clock):
        qn = q.detach().to("cpu", torch.float64).numpy().copy()
        yn = y0.detach().to("cpu", torch.float64).numpy().copy()
        args = (tuple(clock), int(nsteps), int(window_start), int(ckpt_every))
        J = _run_forward(solver, qn, yn, *args)
        ctx.solver = solver
        ctx.inputs = (qn, yn) + args
        ctx.observers = getattr(solver, "observers", None)
        ctx.generation = solver.objective_generation
        ctx.q_meta = (q.device, q.dtype)
        ctx.y_meta = (y0.device, y0.dtype)
        return torch.tensor(J, dtype=torch.float64, device=q.device)

    @staticmethod
    def backward(ctx, gJ):
        s = ctx.solver
        if s.objective_generation != ctx.generation:
            if ctx.observers is not None and getattr(s, "observers", None) is not ctx.observers:
                s.set_observer_points(*ctx.observers)
            _run_forward(s, *ctx.inputs)
        s.objective_backward(float(gJ))
        gq = torch.from_numpy(s.source_gradient()).to(*ctx.q_meta)
        gy = torch.from_numpy(s.state_gradient()).to(*ctx.y_meta)
        return gq, gy, None, None, None, None, None


def observer_energy(solver, q, y0, nsteps, window_start, ckpt_every, clock=(0, 0.0, 1e-6)):
    """J(q, y0) from the initial clock (step, time, dt); differentiable in
    q (block shape) and y0 (5 + block shape)."""
    return ObserverEnergy.apply(q, y0, solver, nsteps, window_start, ckpt_every, clock)


def embed_control(a, shape):
    """Bilinear embedding of a coarse control field a (mj, mi) onto a block
    of `shape` (nj, ni), corners aligned; differentiable."""
    return torch.nn.functional.interpolate(a[None, None], size=tuple(shape), mode="bilinear",
                                           align_corners=True)[0, 0]


def embed_control3(a, shape):
    """Trilinear embedding of a coarse control field a (mk, mj, mi) onto a
    block of `shape` (nk, nj, ni), corners aligned; differentiable."""
    return torch.nn.functional.interpolate(a[None, None], size=tuple(shape), mode="trilinear",
                                           align_corners=True)[0, 0]
