"""``quspin.tools.evolution``: ``expm_multiply_parallel`` (tools/evolution.py:11-16)
and the ODE driver ``evolve`` (tools/evolution.py:144-460).

``evolve`` integrates dy/dt = f(t, y) with scipy's ``ode`` integrators, as the
reference does; the right-hand sides that ``hamiltonian.evolve`` passes in
compute H(t)|v> on the device (operators/hamiltonian_core.py)."""

import numpy as np
from numpy.linalg import norm
from scipy.integrate import ode

from .expm_multiply_parallel_core import (  # noqa: F401
    ExpmMultiplyParallel, expm_multiply_parallel, expm_multiply_sectors)

__all__ = ["expm_multiply_parallel", "ExpmMultiplyParallel", "expm_multiply_sectors", "evolve"]


def evolve(v0, t0, times, f, solver_name="dop853", real=False, stack_state=False, verbose=False, imag_time=False,
           iterate=False, f_params=(), **solver_args):
    """Solve dy/dt = f(t, y, *f_params) from y(t0) = v0 at ``times``
    (tools/evolution.py:144-356): complex states are integrated as float64
    views (``real=False``), or stacked as [Re; Im] (``stack_state``);
    dop853/dopri5 default to nsteps = 2^31-1 and rtol = atol = 1e-9;
    ``imag_time`` restores the initial norm of every column at the output
    times.  Returns an array with time as the last axis, or a generator when
    ``iterate``."""
    v0 = np.asarray(v0)
    ndim = v0.ndim
    if ndim > 2:
        raise ValueError("state mush have ndim < 3.")
    shape0 = v0.shape
    if ndim == 2:
        v0 = v0.ravel()
    if np.iscomplexobj(times):
        raise ValueError("times must be real number(s).")
    n = norm(v0)
    if stack_state:
        if imag_time:
            raise ValueError("imag_time is not compatible with stack_state.")
        complex_valued = False
        v1 = v0.copy()
        v0 = np.zeros(2 * v1.shape[0], dtype=v1.real.dtype)
        v0[:v1.shape[0]] = v1.real
        v0[v1.shape[0]:] = v1.imag
        solver = ode(f)
        solver.set_f_params(*f_params)
    elif real:
        complex_valued = False
        solver = ode(f)
        solver.set_f_params(*f_params)
    else:
        complex_valued = True
        v0 = np.ascontiguousarray(v0, dtype=np.complex128).view(np.float64)
        solver = ode(_cmplx_f)
        solver.set_f_params(f, f_params)
    if solver_name in ["dop853", "dopri5"]:
        if solver_args.get("nsteps") is None:
            solver_args["nsteps"] = np.iinfo(np.int32).max
        if solver_args.get("rtol") is None:
            solver_args["rtol"] = 1e-9
        if solver_args.get("atol") is None:
            solver_args["atol"] = 1e-9
    solver.set_integrator(solver_name, **solver_args)
    solver.set_initial_value(v0, t0)
    fmt = (complex_valued, stack_state, imag_time, n, shape0)
    if np.isscalar(times):
        if times == t0:
            return _format_output(v0, *fmt)
        return _integrate(solver, times, fmt)
    if iterate:
        return _evolve_iter(solver, v0, t0, times, verbose, fmt)
    out = list(_evolve_iter(solver, v0, t0, times, verbose, fmt))
    dt = np.complex128 if (complex_valued or stack_state) else np.float64
    v = np.stack(out, axis=0).astype(dt)
    return v.transpose() if v.ndim == 2 else v.transpose((1, 2, 0))


def _cmplx_f(t, y, f, f_params):
    return f(t, y.view(np.complex128), *f_params).view(np.float64)


def _format_output(y, complex_valued, stack_state, imag_time, n, shape0):
    if stack_state:
        half = y.size // 2
        yout = y[:half].astype(np.complex128).reshape(shape0)
        yout[...] += 1j * y[half:].reshape(shape0)
    elif complex_valued:
        yout = y.view(np.complex128).reshape(shape0).copy()
    else:
        yout = y.reshape(shape0).copy()
    if imag_time:
        yout /= norm(yout, axis=0) / n
    return yout


def _integrate(solver, t, fmt):
    solver.integrate(t)
    if solver.successful():
        return _format_output(solver._y, *fmt)
    raise RuntimeError("failed to evolve to time {0}, nsteps might be too small".format(t))


def _evolve_iter(solver, v0, t0, times, verbose, fmt):
    for t in times:
        y = _format_output(v0, *fmt) if t == t0 else _integrate(solver, t, fmt)
        if verbose:
            print("evolved to time {0}, norm of state(s) {1}".format(t, norm(y, axis=0)))
        yield y
