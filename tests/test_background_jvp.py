"""Derivatives of the background tables (coffe_background_jvp, background.c) with
respect to (Omega_cdm, Omega_baryon, Omega_gamma, w0, wa, h):
- the values equal coffe_background_compute's bit for bit (one code path);
- closed forms in h: Hz, conformal_Hz scale as h, conformal_Hz_prime as h^2 and chi as
  1/h, while D1, f, G1 and G2 do not depend on h;
- every direction against central differences of the tables (step 1e-5 max(|x|, 0.1)),
  to 1e-7 of each table's largest derivative (plus 1e-9 of the table per unit of
  the parameter, the noise of the difference);
- linearity in the direction (1e-12)."""
import numpy as np
import pytest

from coffe_amd import background as B

KW = dict(omega_cdm=0.26, omega_baryon=0.049, omega_gamma=9e-5, w0=-0.95, wa=0.1, h=0.68, bins=600,
          s1=np.full(11, 0.3), s2=np.full(11, 0.2), fevo1=np.full(11, 0.5), fevo2=np.full(11, 0.1))


def test_values_match_compute():
    tab, _ = B.jvp({"omega_cdm": 1.0}, **KW)
    ref = B.compute(**KW)
    for n in B.NAMES:
        assert np.array_equal(tab[n], ref[n]), n


def test_closed_forms_in_h():
    tab, der = B.jvp({"h": 1.0}, **KW)
    h = KW["h"]
    for n in ("Hz", "conformal_Hz"):
        np.testing.assert_allclose(der[n], tab[n] / h, rtol=1e-14, atol=0)
    np.testing.assert_allclose(der["conformal_Hz_prime"], 2 * tab["conformal_Hz_prime"] / h, rtol=1e-14, atol=0)
    np.testing.assert_allclose(der["chi"], -tab["chi"] / h, rtol=1e-13, atol=0)
    for n in ("D1", "f", "z", "a"):
        assert np.all(der[n] == 0), n
    for n in ("G1", "G2"):
        assert np.max(np.abs(der[n])) <= 1e-12 * np.max(np.abs(tab[n])), n


@pytest.mark.parametrize("param", B.PARAMS)
def test_against_central_differences(param):
    tab, der = B.jvp({param: 1.0}, **KW)
    x0 = KW[param]
    step = 1e-5 * max(abs(x0), 0.1)     # above the rounding noise of the RK4 tables
    up = B.compute(**{**KW, param: x0 + step})
    dn = B.compute(**{**KW, param: x0 - step})
    for n in B.NAMES:
        fd = (up[n] - dn[n]) / (2 * step)
        # a floor for the tables that do not depend on the parameter (G1, G2 in h): the
        # difference quotient is then rounding noise of the table over the step
        floor = 1e-9 * np.max(np.abs(tab[n])) / max(abs(x0), 0.1)
        err = np.max(np.abs(der[n] - fd))
        assert err <= 1e-7 * np.max(np.abs(der[n])) + floor, (param, n, err)


def test_linear_in_direction():
    d1 = {"omega_cdm": 0.7, "w0": -0.3}
    d2 = {"omega_baryon": 1.1, "wa": 0.4, "h": 2.0}
    _, a = B.jvp(d1, **KW)
    _, b = B.jvp(d2, **KW)
    _, c = B.jvp({**d1, **d2}, **KW)
    for n in B.NAMES:
        np.testing.assert_allclose(c[n], a[n] + b[n], rtol=0, atol=1e-12 * max(np.max(np.abs(c[n])), 1e-300))


def test_unknown_parameter():
    with pytest.raises(ValueError):
        B.jvp({"omega_m": 1.0}, **KW)
