"""The drop-in Coffe object (coffe_amd.coffe) on the CPU: parameter validation and
error behaviour as the reference's own tests exercise them (test_coffe.py:47-120),
defaults (parser.c:733-1033) and the result records."""
import numpy as np
import pytest

from coffe_amd.coffe import Coffe
from coffe_amd.representation import Covariance, Multipoles


def test_defaults():
    c = Coffe()
    assert c.omega_m == pytest.approx(0.3) and c.h == 0.67 and c.omega_baryon == 0.05
    assert list(c.l) == [0, 2, 4] and list(c.z_mean) == [1.0]
    assert len(c.sep) == 100 and c.sep[0] == 5.0
    assert c.has_density and not c.has_rsd and not c.has_d1 and not c.has_lensing
    assert c.integrals_fftlog_sampling == 10000 and c.covariance_integration_sampling == 8000


def test_bias():
    c = Coffe()
    with pytest.raises(ValueError):
        x = np.linspace(-1, 10, 100)
        c.set_galaxy_bias1(x, np.zeros(len(x)))
    with pytest.raises(ValueError):
        x = np.linspace(0, 100, 100)
        c.set_galaxy_bias1(x, np.zeros(len(x)))
    with pytest.raises(ValueError):
        x = np.linspace(0, 1, 100)
        c.set_galaxy_bias1(x, np.zeros(len(x)))
        c.galaxy_bias1(2)
    with pytest.raises(ValueError):
        c.set_galaxy_bias2([0, 2, 1, 3, 4], np.ones(5))
    with pytest.raises(ValueError):
        c.set_galaxy_bias2(np.linspace(0, 5, 6), np.ones(5))


def test_parameters():
    c = Coffe()
    with pytest.raises(ValueError):
        c.omega_cdm = 1.1
    with pytest.raises(ValueError):
        c.h = -1
    with pytest.raises(ValueError):   # omega_baryon = 0.05 by default
        c.omega_cdm = 0.99
    with pytest.raises(ValueError):
        c.omega_m = 0.99999999
    with pytest.raises(ValueError):
        c.z_mean = (-1,)
    with pytest.raises(TypeError):
        c.z_mean = 0.1
    with pytest.raises(TypeError):
        c.l = [0.5]
    with pytest.raises(ValueError):
        c.sep = [0.0, 10.0]
    with pytest.raises(ValueError):   # only the default tables exist (coffe_find_integral)
        c.integral(r=100, n=4, l=4)
    with pytest.raises(TypeError):
        c.integral(r=100, n=0.5, l=0)
    with pytest.raises(AttributeError):
        c.set_parameters(no_such_parameter=1)
    c.has_g4 = True
    c.has_g5 = False
    assert c.has_g4 and not c.has_g5 and c.parameters["has_g4"]
    c.reset_contributions()
    assert not any(v for k, v in c.parameters.items() if k.startswith("has_"))


def test_lensing_plan():
    """Which integrated terms go where (signal.c:556-771): the flat-sky forms with the
    flat-sky switches, the rest to the full-sky integrals; the cosmological tangent
    refuses the integrated terms."""
    c = Coffe(has_lensing=True, has_density=False)
    c.set_power_spectrum_linear(np.logspace(-4, 1, 50), np.logspace(-4, 1, 50) ** -1)
    assert c._lensing_plan() == dict(flat_ll=False, flat_dl=False, flags={"len"})
    c.has_flatsky_nonlocal = True
    assert c._lensing_plan() == dict(flat_ll=True, flat_dl=False, flags=None)
    c.has_density = True
    plan = c._lensing_plan()
    assert plan["flat_ll"] and not plan["flat_dl"] and plan["flags"] == {"den", "len", "flat_ll"}
    c.has_flatsky_local_nonlocal = True
    plan = c._lensing_plan()
    assert plan["flat_ll"] and plan["flat_dl"] and plan["flags"] is None
    c.has_d2 = True          # d2-lensing stays full sky
    assert c._lensing_plan()["flags"] == {"den", "d2", "len", "flat_ll", "flat_dl"}
    c.has_d2 = False
    c.has_only_cross_correlations = True
    plan = c._lensing_plan()
    assert not plan["flat_ll"] and plan["flat_dl"]
    c.reset_contributions()
    c.has_g4 = True
    assert c._lensing_plan() == dict(flat_ll=False, flat_dl=False, flags={"g4"})
    with pytest.raises(NotImplementedError):
        c.multipoles_cosmo_jvp({"h": 1.0})


def test_coordinates_and_contributions_checked_first():
    c = Coffe()
    c.z_mean = []
    with pytest.raises(ValueError):
        c.compute_multipoles_bulk()
    with pytest.raises(ValueError):
        c.compute_covariance_bulk()
    c = Coffe()
    c.reset_contributions()
    with pytest.raises(ValueError):
        c.compute_multipole(z=0.1, r=100, l=2)
    c = Coffe()
    with pytest.raises(ValueError):   # no P(k) yet
        c.integral(r=100, n=0, l=0)


def test_power_spectrum_input_checks():
    c = Coffe()
    with pytest.raises(ValueError):
        c.set_power_spectrum_linear([3, 2, 1, 4, 5], [1, 1, 1, 1, 1])
    with pytest.raises(ValueError):
        c.set_power_spectrum_linear([1, 2, 3, 4, 5], [1, 1, 1, 1])
    c.set_power_spectrum_linear(np.logspace(-4, 1, 10), np.ones(10))
    assert c.k_min == 1e-4 and c.k_max == 10.0


def test_records():
    m = Multipoles(z=1.0, r=10.0, l=2, value=0.5)
    assert m.to_dict() == {"z": 1.0, "r": 10.0, "l": 2, "value": 0.5}
    c = Covariance(r1=1.0, r2=2.0, l1=0, l2=2, z=1.0, value=3.0)
    assert c.l2 == 2 and c.to_dict()["value"] == 3.0
    with pytest.raises(TypeError):
        Multipoles(z=1.0, r=1.0, l=0)


def test_average_multipoles_coordinates_checked():
    from coffe_amd.coffe import Coffe
    c = Coffe()
    c.z_min = [0.9, 1.2]
    c.z_max = [1.1]
    with pytest.raises(ValueError):
        c.compute_average_multipoles_bulk()
    c.z_max = [1.1, 1.1]
    with pytest.raises(ValueError):
        c.compute_average_multipoles_bulk()
    with pytest.raises(ValueError):
        c.z_min = [12.0]


def test_covariance_populations_and_densities():
    """covariance_populations (coffe.pyx:2160-2192: default 1, 1, 1, 1, length 4, values
    in 1..4), number_density1/2 (coffe.pyx:2097-2157: values in [0, 10)) and
    has_binned_covariance (coffe.pyx:2525-2545)."""
    c = Coffe()
    assert list(c.covariance_populations) == [1, 1, 1, 1]
    c.covariance_populations = [1, 2, 1, 2]
    assert list(c.covariance_populations) == [1, 2, 1, 2]
    with pytest.raises(ValueError):
        c.covariance_populations = [1, 2, 1]
    with pytest.raises(ValueError):
        c.covariance_populations = [1, 2, 5, 1]
    with pytest.raises(TypeError):
        c.covariance_populations = 3
    with pytest.raises(TypeError):
        c.covariance_populations = [1, None, 1, 1]
    assert list(c.covariance_populations) == [1, 2, 1, 2]
    c.number_density2 = [2e-4]
    assert list(c.number_density2) == [2e-4]
    with pytest.raises(ValueError):
        c.number_density2 = [11.0]
    with pytest.raises(ValueError):
        c.number_density1 = [-1e-3]
    with pytest.raises(TypeError):
        c.number_density1 = 1e-3
    assert c.has_binned_covariance is False
    c.has_binned_covariance = 1
    assert c.has_binned_covariance is True


def test_covariance_lengths_checked():
    """Mismatching per-bin covariance lists raise ValueError before any computation
    (coffe.pyx:3448-3452)."""
    c = Coffe(z_mean=[1.0, 1.5], deltaz=[0.1, 0.1], fsky=[0.2, 0.2], pixelsize=[10.0, 10.0],
              number_density1=[1e-3, 1e-3], number_density2=[1e-3])
    c.set_power_spectrum_linear(np.logspace(-4, 1, 64), np.ones(64))
    with pytest.raises(ValueError):
        c.compute_covariance_bulk()


def test_flat_weights_of_background_rejects_unsupported_weights():
    """Only flat-sky density + RSD weights with a constant bias depend on the background
    through D1^2 and f alone; other LocalWeights are refused (no silent wrong gradient)."""
    import torch
    from coffe_amd.autograd import LocalWeights, flat_weights_of_background
    theta = torch.zeros(6, dtype=torch.float64)
    for flatsky, d1, const in [(False, False, True), (True, True, True), (True, False, False)]:
        lw = LocalWeights.__new__(LocalWeights)
        lw.f_override, lw.flatsky, lw.d1, lw.constant_bias, lw.groups = True, flatsky, d1, const, {}
        with pytest.raises(ValueError):
            flat_weights_of_background(lw, theta, theta, [1.0], 1.5)
