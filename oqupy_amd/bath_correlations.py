"""Environment auto-correlations and the eta-function integrals.

The influence functional of the process-tensor path is fed by
``correlation_2d_integral``; this module restates the reference's
``oqupy/bath_correlations.py`` (classes at 29-278 and 331-759) with the same
integrands, intervals, tolerances and sign conventions, so that the influence
tensors handed to the MI355X engine are those the reference would build. The
quadrature stays on the host in FP64 (scipy QUADPACK), as in the reference.
"""

from functools import lru_cache
from typing import Callable, Optional, Text

import numpy as np
from scipy import integrate

from oqupy_amd.base_api import BaseAPIClass
from oqupy_amd.config import INTEGRATE_EPSREL, SUBDIV_LIMIT
from oqupy_amd.util import check_true


class BaseCorrelations(BaseAPIClass):
    """Base class of environment auto-correlations
    (oqupy/bath_correlations.py:29-118)."""

    def correlation(self, tau, epsrel: Optional[float] = INTEGRATE_EPSREL,
                    subdiv_limit: Optional[int] = SUBDIV_LIMIT):
        """Auto-correlation ``C(tau) = <F(t) F(t-tau)>``."""
        raise NotImplementedError(
            f"{type(self).__name__} has no correlation implementation.")

    def correlation_2d_integral(self, delta: float, time_1: float,
                                time_2: Optional[float] = None,
                                shape: Optional[Text] = 'square',
                                epsrel: Optional[float] = INTEGRATE_EPSREL,
                                subdiv_limit: Optional[int] = SUBDIV_LIMIT
                                ) -> complex:
        """Double integral of ``C`` over a square, upper triangle or
        rectangle of the (t', t'') plane."""
        raise NotImplementedError(
            f"{type(self).__name__} has no correlation_2d_integral "
            "implementation.")


class CustomCorrelations(BaseCorrelations):
    """A user-given correlation function ``C(tau)``
    (oqupy/bath_correlations.py:120-278).

    The 2D integrals are done with ``scipy.integrate.dblquad`` on the real and
    imaginary part, as in the reference (lines 253-278).
    """

    def __init__(self, correlation_function: Callable[[float], complex],
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        try:
            func = np.vectorize(correlation_function)
            complex(func(1.0))
        except Exception as err:
            raise AssertionError("Correlation function must be vectorizable "
                                 "and must return float.") from err
        self.correlation_function = func
        super().__init__(name, description)

    def correlation(self, tau, epsrel: Optional[float] = None,
                    subdiv_limit: Optional[int] = None):
        """``C(tau)`` as given."""
        return self.correlation_function(tau)

    @lru_cache(maxsize=2 ** 10, typed=False)
    def correlation_2d_integral(self, delta: float, time_1: float,
                                time_2: Optional[float] = None,
                                shape: Optional[Text] = 'square',
                                epsrel: Optional[float] = INTEGRATE_EPSREL,
                                subdiv_limit: Optional[int] = SUBDIV_LIMIT
                                ) -> complex:
        """``int_{t1}^{t2} dt' int_{lo(t')}^{hi(t')} dt'' C(t'-t'')`` with
        ``t2 = t1 + delta`` unless ``shape == 'rectangle'``."""
        if time_2 is None:
            time_2 = time_1 + delta
        else:
            assert shape == 'rectangle', (
                "parameter 'time_2' can only be used in conjunction with "
                "'shape' = ``rectangle`` !")
        upper = {'square': lambda x: delta,
                 'upper-triangle': lambda x: x - time_1,
                 'rectangle': lambda x: delta}[shape]
        parts = []
        for take in (np.real, np.imag):
            func = lambda y, x, take=take: take(self.correlation(x - y))
            parts.append(integrate.dblquad(func=func, a=time_1, b=time_2,
                                           gfun=lambda x: 0.0, hfun=upper,
                                           epsrel=epsrel)[0])
        return parts[0] + 1.0j * parts[1]


def _hard_cutoff(omega, omega_c):
    return np.heaviside(omega_c - omega, 0)


def _exponential_cutoff(omega, omega_c):
    return np.exp(-omega / omega_c)


def _gaussian_cutoff(omega, omega_c):
    return np.exp(-(omega / omega_c) ** 2)


# cutoff name -> X(omega, omega_c)  (oqupy/bath_correlations.py:285-305)
CUTOFF_DICT = {
    'hard': _hard_cutoff,
    'exponential': _exponential_cutoff,
    'gaussian': _gaussian_cutoff,
}


def _complex_quad(integrand, a, b, epsrel, limit) -> complex:
    """Real and imaginary parts integrated separately with QUADPACK
    (oqupy/bath_correlations.py:312-329)."""
    re = integrate.quad(lambda x: np.real(integrand(x)), a=a, b=b,
                        epsrel=epsrel, limit=limit)[0]
    im = integrate.quad(lambda x: np.imag(integrand(x)), a=a, b=b,
                        epsrel=epsrel, limit=limit)[0]
    return re + 1j * im


class CustomSD(BaseCorrelations):
    r"""Correlations of a thermal bath with spectral density
    ``J(w) = j(w) X(w, w_c)`` (oqupy/bath_correlations.py:331-664)."""

    def __init__(self, j_function: Callable[[float], float], cutoff: float,
                 cutoff_type: Optional[Text] = 'exponential',
                 temperature: Optional[float] = 0.0,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        try:
            func = np.vectorize(j_function)
            float(func(1.0))
        except Exception as err:
            raise AssertionError("Spectral density must be vectorizable "
                                 "and must return float.") from err
        self.j_function = func
        try:
            self.cutoff = float(cutoff)
        except Exception as err:
            raise AssertionError("Cutoff must be a float.") from err
        assert cutoff_type in CUTOFF_DICT, \
            f"Cutoff type must be one of: {CUTOFF_DICT.keys()}"
        self.cutoff_type = cutoff_type
        try:
            temp = float(temperature)
        except Exception as err:
            raise AssertionError("Temperature must be a float.") from err
        if temp < 0.0:
            raise ValueError(f"Temperature must be >= 0.0 (but is {temp})")
        self.temperature = temp
        super().__init__(name, description)

    def _cutoff_function(self, omega):
        return CUTOFF_DICT[self.cutoff_type](omega, self.cutoff)

    def spectral_density(self, omega):
        """``J(w)`` including the cutoff."""
        return self.j_function(omega) * self._cutoff_function(omega)

    def _integrate_omega(self, integrand, epsrel, subdiv_limit) -> complex:
        """Integral over [0, w_c], plus [w_c, inf) unless the cutoff is hard
        (oqupy/bath_correlations.py:502-512, 577-587)."""
        total = _complex_quad(integrand, 0.0, self.cutoff, epsrel,
                              subdiv_limit)
        if self.cutoff_type != "hard":
            total += _complex_quad(integrand, self.cutoff, np.inf, epsrel,
                                   subdiv_limit)
        return total

    def correlation(self, tau, epsrel: Optional[float] = INTEGRATE_EPSREL,
                    subdiv_limit: Optional[int] = SUBDIV_LIMIT,
                    matsubara: Optional[bool] = False):
        r"""``C(tau) = int J(w) [cos(w tau) coth(w/2T) - i sin(w tau)] dw``
        (oqupy/bath_correlations.py:446-516)."""
        if matsubara:
            tau = -1j * tau
        temp = self.temperature
        if temp == 0.0:
            check_true(matsubara is False,
                       'Matsubara correlations only defined for '
                       'temperature > 0')

            def integrand(w):
                return self.spectral_density(w) * np.exp(-1j * w * tau)
        else:
            def integrand(w):
                boltz = np.exp(-w / temp)
                if boltz > np.finfo(float).eps:
                    return (self.spectral_density(w)
                            * (np.exp(-1j * tau * w)
                               + np.exp(-(w / temp - 1j * tau * w)))
                            / (1 - boltz))
                return self.spectral_density(w) * np.exp(-1j * w * tau)
        total = self._integrate_omega(integrand, epsrel, subdiv_limit)
        return total.real if matsubara else total

    @lru_cache(maxsize=2 ** 10, typed=False)
    def eta_function(self, tau, epsrel: Optional[float] = INTEGRATE_EPSREL,
                     subdiv_limit: Optional[int] = SUBDIV_LIMIT,
                     matsubara: Optional[bool] = False):
        r"""The twice-integrated correlation
        ``eta(tau) = int_0^tau int_0^t' C(t'') dt'' dt'``, computed as
        ``-int J(w)/w^2 [...] dw`` with the integrands of
        ``oqupy/bath_correlations.py:557-591``."""
        if matsubara:
            tau = -1j * tau
        temp = self.temperature
        if temp == 0.0:
            check_true(matsubara is False,
                       'Matsubara correlations only defined for '
                       'temperature > 0')

            def integrand(w):
                return self.spectral_density(w) / w ** 2 * (
                    (np.exp(-1j * w * tau) - 1) + 1j * w * tau)
        else:
            def integrand(w):
                boltz = np.exp(-w / temp)
                if boltz > np.finfo(float).eps:
                    return (self.spectral_density(w) / w ** 2
                            * (((np.exp(-1j * tau * w)
                                 + np.exp(-(w / temp - 1j * tau * w)))
                                - boltz - 1) / (1 - boltz)
                               + 1j * tau * w))
                return (self.spectral_density(w) / w ** 2
                        * (np.exp(-1j * w * tau) - 1 + 1j * w * tau))
        total = self._integrate_omega(integrand, epsrel, subdiv_limit)
        if matsubara:
            total = total.real
        return -total

    def correlation_2d_integral(self, delta: float, time_1: float,
                                time_2: Optional[float] = None,
                                shape: Optional[Text] = 'square',
                                epsrel: Optional[float] = INTEGRATE_EPSREL,
                                subdiv_limit: Optional[int] = SUBDIV_LIMIT,
                                matsubara: Optional[bool] = False) -> complex:
        """2D correlation integrals as finite differences of ``eta``
        (oqupy/bath_correlations.py:593-664)."""
        kw = {'epsrel': epsrel, 'subdiv_limit': subdiv_limit,
              'matsubara': matsubara}
        eta = self.eta_function
        if shape == 'upper-triangle':
            value = eta(time_1 + delta, **kw) - eta(time_1, **kw)
        elif shape == 'square':
            value = (eta(time_1 + delta, **kw) - 2.0 * eta(time_1, **kw)
                     + eta(time_1 - delta, **kw))
        elif shape == 'rectangle':
            value = (eta(time_2, **kw) - eta(time_1, **kw)
                     - eta(time_2 - delta, **kw) + eta(time_1 - delta, **kw))
        else:
            raise NotImplementedError(f"Shape '{shape}' not implemented.")
        return value.real if matsubara else value


class PowerLawSD(CustomSD):
    r"""``J(w) = 2 alpha w^zeta / w_c^(zeta-1) X(w, w_c)``
    (oqupy/bath_correlations.py:667-759)."""

    def __init__(self, alpha: float, zeta: float, cutoff: float,
                 cutoff_type: Text = 'exponential',
                 temperature: Optional[float] = 0.0,
                 name: Optional[Text] = None,
                 description: Optional[Text] = None) -> None:
        try:
            self.alpha = float(alpha)
        except Exception as err:
            raise AssertionError("Alpha must be a float.") from err
        try:
            self.zeta = float(zeta)
        except Exception as err:
            raise AssertionError("Nu must be a float.") from err
        try:
            wc = float(cutoff)
        except Exception as err:
            raise AssertionError("Cutoff must be a float.") from err
        a, z = self.alpha, self.zeta
        super().__init__(lambda w: 2.0 * a * w ** z * wc ** (1 - z),
                         cutoff=wc, cutoff_type=cutoff_type,
                         temperature=temperature, name=name,
                         description=description)
