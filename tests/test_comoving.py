"""Nyx::integrate_comoving_a (Source/Driver/comoving.cpp:630-689) against closed
forms.  The function is host arithmetic in the library (no GPU needed):
  Einstein-de Sitter (OmM = 1):  a^(3/2) = a0^(3/2) + (3/2) H0 dt
  Lambda only (OmM = 0):         a = a0 exp(H0 dt)
  radiation only (OmR = 1):      a^2 = a0^2 + 2 H0 dt
  comoving_type <= 0:            a = a0 + h dt   (the reference's linear branch)
with H0 = 100 h.  The RK2 step doubling stops at 1e-8 relative change, so the
tolerance is a few 1e-8."""
import math

import pytest

import nyx_amd


@pytest.mark.parametrize("a0", [1 / 101, 0.1, 0.5])
@pytest.mark.parametrize("growth", [0.01, 0.05, 0.1])
def test_einstein_de_sitter(a0, growth):
    # steps that change a by 1-10% (Nyx limits it to relative_max_change_a = 1%)
    h = 0.675
    H0 = 100 * h
    dt = ((a0 * (1 + growth)) ** 1.5 - a0 ** 1.5) / (1.5 * H0)
    a = nyx_amd.integrate_comoving_a(a0, dt, 1.0, 0.0, h, 1)
    exact = (a0 ** 1.5 + 1.5 * H0 * dt) ** (2 / 3)
    assert abs(a / exact - 1) < 5e-8


def test_one_large_step_is_bounded():
    # a 22-fold growth in one call: the reference's 11 doubling passes (1024
    # substeps) end before the 1e-8 criterion; the result is still within 1e-6
    h = 0.675
    a0, dt = 1 / 101, 1e-3
    a = nyx_amd.integrate_comoving_a(a0, dt, 1.0, 0.0, h, 1)
    exact = (a0 ** 1.5 + 1.5 * 100 * h * dt) ** (2 / 3)
    assert abs(a / exact - 1) < 1e-6


def test_lambda_and_radiation():
    h, a0, dt = 0.7, 0.3, 5e-4
    H0 = 100 * h
    a = nyx_amd.integrate_comoving_a(a0, dt, 0.0, 0.0, h, 1)
    assert abs(a / (a0 * math.exp(H0 * dt)) - 1) < 5e-8
    a = nyx_amd.integrate_comoving_a(a0, dt, 0.0, 1.0, h, 1)
    assert abs(a / math.sqrt(a0 ** 2 + 2 * H0 * dt) - 1) < 5e-8


def test_linear_branch_and_static():
    assert nyx_amd.integrate_comoving_a(0.2, 0.5, 0.3, 0.0, 0.7, 0) == 0.2 + 0.7 * 0.5
    assert nyx_amd.integrate_comoving_a(0.2, 0.5, 0.3, 0.0, 0.0, 1) == 0.2  # h = 0: a is constant
    with pytest.raises(nyx_amd.NyxPMError):
        nyx_amd.integrate_comoving_a(0.0, 0.5, 0.3, 0.0, 0.7, 1)
