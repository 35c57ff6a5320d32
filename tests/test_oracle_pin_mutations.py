"""The closed-form pins of tests/test_oracle_pins.py catch plausible
mistakes in the oracle: each mistake is planted in the oracle on the CPU (a
wrapper around the oracle function, installed with monkeypatch) and the pin
that fixes it must then fail."""
import numpy as np
import pytest

from oracle import fd as ofd
from oracle import newton_krylov as onk
from tests import test_oracle_pins as pins


def _fails(fn, *args):
    with pytest.raises(AssertionError):
        fn(*args)


def test_count_off_by_one_fails_the_count_pins(monkeypatch):
    real = onk.gmres

    def plus_one(*a, **k):  # one iteration too many (e.g. the initial residual counted)
        x, it, conv, ratio = real(*a, **k)
        return x, it + 1, conv, ratio

    monkeypatch.setattr(onk, "gmres", plus_one)
    _fails(pins.test_gmres_one_iteration_for_the_identity_and_an_exact_inverse)
    _fails(pins.test_gmres_counts_the_krylov_dimension, 3)
    _fails(pins.test_gmres_counts_only_the_eigenvalues_b_touches)
    _fails(pins.test_gmres_zero_rhs_takes_no_iteration)


def test_count_per_cycle_fails_the_restart_pin(monkeypatch):
    real = onk.gmres

    def per_cycle(A, b, M=None, rtol=1e-8, restart=30, maxiter=50):  # counts restart cycles, not Arnoldi steps
        x, it, conv, ratio = real(A, b, M=M, rtol=rtol, restart=restart, maxiter=maxiter)
        return x, -(-it // restart), conv, ratio

    monkeypatch.setattr(onk, "gmres", per_cycle)
    _fails(pins.test_gmres_restart_cycles_count_every_arnoldi_step)


def test_upwind_flipped_at_zero_velocity_fails_the_direction_pin(monkeypatch):
    real = ofd.d1_upwind_2d

    def forward_at_zero(f, vel, dy, dx, axis):  # v = 0 taken as negative (forward quotient)
        v = np.where(np.asarray(vel) == 0.0, -1.0, vel)
        return real(f, v if np.ndim(vel) else float(v), dy, dx, axis)

    monkeypatch.setattr(ofd, "d1_upwind_2d", forward_at_zero)
    _fails(pins.test_upwind_quadratic_picks_backward_for_nonnegative_velocity)


def test_upwind_direction_swapped_fails_the_direction_pin(monkeypatch):
    real = ofd.d1_upwind_2d

    def swapped(f, vel, dy, dx, axis):  # backward where v < 0, forward where v > 0
        v = -np.asarray(vel, dtype=float)
        v = np.where(v == 0.0, 0.0, v)
        return real(f, v if np.ndim(vel) else float(v), dy, dx, axis)

    monkeypatch.setattr(ofd, "d1_upwind_2d", swapped)
    _fails(pins.test_upwind_quadratic_picks_backward_for_nonnegative_velocity)
