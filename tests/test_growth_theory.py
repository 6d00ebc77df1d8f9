"""The linear theory behind the growth-rate checks, pinned on the CPU by its
known limits: the cold two-beam filamentation rate reduces to the textbook
non-relativistic Weibel rate beta K / sqrt(K^2 + 1) for slow beams, tends to
beta / sqrt(gamma) for short wavelengths and to K beta gamma for long ones."""
import math

import pytest

from tests.test_gpu_physics import filamentation_rate


@pytest.mark.parametrize("K", [0.3, 1.0, 1.65, 4.0, 20.0])
def test_slow_beams_reduce_to_the_weibel_rate(K):
    beta = 1e-3
    assert filamentation_rate(K, beta) == pytest.approx(beta * K / math.sqrt(K * K + 1.0), rel=1e-5)


@pytest.mark.parametrize("beta", [0.1, 0.3, 0.6, 0.9])
def test_limits_in_wavenumber(beta):
    gamma = 1.0 / math.sqrt(1.0 - beta * beta)
    assert filamentation_rate(1e4, beta) == pytest.approx(beta / math.sqrt(gamma), rel=1e-6)
    assert filamentation_rate(1e-5, beta) == pytest.approx(1e-5 * beta * gamma, rel=1e-6)


@pytest.mark.parametrize("K,beta", [(1.65, 0.3), (3.3, 0.3), (0.5, 0.8)])
def test_rate_is_a_root_and_grows_with_k(K, beta):
    gamma = 1.0 / math.sqrt(1.0 - beta * beta)
    G = filamentation_rate(K, beta)
    resid = G ** 4 + G * G * (K * K + gamma ** -3) - K * K * beta * beta / gamma
    assert abs(resid) <= 1e-12
    assert filamentation_rate(1.1 * K, beta) > G
