"""Settings files and the parameter bookkeeping of the Coffe object, on the CPU
(tests/test_coffe.py:21-47 of the reference: from_file of the benchmark settings with a
DeprecationWarning, to_file -> from_file round trip of the parameters)."""
import os

import numpy as np
import pytest

from coffe_amd.coffe import Coffe

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _plain(p):
    return {k: list(v) if isinstance(v, (list, np.ndarray)) else v for k, v in p.items()}


def test_from_file_benchmark_settings():
    """The reference's benchmark settings file (legacy option names throughout)."""
    with pytest.warns(DeprecationWarning):
        c = Coffe.from_file(os.path.join(GOLD, "settings_benchmark.cfg"))
    # the current names override the legacy ones read first (coffe.pyx:566-570)
    assert list(c.z_mean) == [1.0] and list(c.deltaz) == [0.2] and list(c.z_min) == [0.9]
    assert list(c.z_max) == [1.1] and list(c.mu) == [0.7]
    # covariance_density feeds both number densities; covariance_fsky / pixelsize map over
    assert list(c.number_density1) == [0.1, 0.2, 0.01] and list(c.number_density2) == [0.1, 0.2, 0.01]
    assert list(c.fsky) == [0.3, 0.5, 0.2] and list(c.pixelsize) == [10.0, 10.0, 10.0]
    assert list(c.l) == [0, 2, 4]
    assert c.has_density and not c.has_rsd and not c.has_binned_covariance and not c.has_only_cross_correlations
    assert c.omega_m == pytest.approx(0.3) and c.omega_baryon == 0.05 and c.omega_gamma == 9e-5
    assert c.h == 0.67 and c.w0 == -1.0 and c.wa == 0.0
    assert c.k_min == 1e-5 and c.k_max == 300.0 and c.sigma8 == 0.8156 and c.n_s == 0.96
    assert c.background_sampling == 10000 and c.integration_sampling == 750000
    assert c.pk_type == "linear" and c.covariance_integration_method == "standard"
    assert c.covariance_interpolation_method == "bicubic"
    # input_galaxy_bias1 present: a constant bias of 'galaxy_bias1' on 100 points in [0, 10]
    assert np.allclose(c.galaxy_bias1_redshifts(), np.linspace(0, 10, 100))
    assert np.allclose(c.galaxy_bias1_values(), 1.0) and np.allclose(c.magnification_bias2_values(), 0.0)


def test_to_from_file_defaults(tmp_path):
    out = Coffe()
    out.to_file(tmp_path / "test.cfg")
    back = Coffe.from_file(tmp_path / "test.cfg")
    assert _plain(out.parameters) == _plain(back.parameters)


def test_to_from_file_changed(tmp_path):
    out = Coffe(omega_m=0.31, omega_baryon=0.048, h=0.7, w0=-0.9, wa=0.1, has_rsd=True, has_d1=True,
                sep=[10.0, 20.5, 33.25], l=[0, 2], z_mean=[0.5, 1.5], deltaz=[0.1, 0.3], mu=[0.1, -0.4],
                covariance_populations=[1, 2, 1, 2], number_density1=[1e-3, 2e-3], number_density2=[3e-3, 4e-3],
                has_only_cross_correlations=True, covariance_integration_method="fftlog",
                covariance_integration_sampling=4096, integrals_fftlog_sampling=8192, multipole_quadrature_nodes=64,
                sigma8=0.81, N_ncdm=0, l_max_g=10, k_min=1e-4, k_max=100.0)
    zb = np.linspace(0, 5, 11)
    out.set_galaxy_bias1(zb, 1 + zb)
    out.set_magnification_bias2(zb, 0.3 * zb)
    out.to_file(tmp_path / "test.cfg")
    back = Coffe.from_file(tmp_path / "test.cfg")
    assert _plain(out.parameters) == _plain(back.parameters)
    assert back.omega_m == pytest.approx(0.31) and back.omega_cdm == pytest.approx(0.31 - 0.048)
    assert np.allclose(back.galaxy_bias1_redshifts(), zb) and np.allclose(back.galaxy_bias1_values(), 1 + zb)
    assert np.allclose(back.magnification_bias2_values(), 0.3 * zb)
    text = (tmp_path / "test.cfg").read_text()
    assert "has_rsd = 1\n" in text and "sep = [10.0, 20.5, 33.25]\n" in text and "pk_type = linear\n" in text


def test_legacy_contributions_and_ignored_options(tmp_path):
    cfg = tmp_path / "legacy.cfg"
    cfg.write_text('correlation_contributions = ["den", "rsd", "d1"];\nmultipoles = [0, 2];\n'
                   "covariance_step_size = 10.0;\nonly_cross_correlations = 1;\ncovariance_window = 1;\n")
    with pytest.warns(DeprecationWarning) as record:
        c = Coffe.from_file(cfg)
    assert c.has_density and c.has_rsd and c.has_d1 and not c.has_d2
    assert list(c.l) == [0, 2] and c.has_only_cross_correlations and c.has_binned_covariance
    assert any("covariance_step_size" in str(w.message) and "ignored" in str(w.message) for w in record)


def test_integrated_contributions_round_trip(tmp_path):
    """The lensing and potential contributions of a settings file (parser.c's
    correlation_contributions names len, g4, g5) survive to_file / from_file."""
    cfg = tmp_path / "intg.cfg"
    cfg.write_text('correlation_contributions = ["den", "len", "g4", "g5"];\nflatsky_nonlocal = 1;\n')
    with pytest.warns(DeprecationWarning):
        c = Coffe.from_file(cfg)
    assert c.has_density and c.has_lensing and c.has_g4 and c.has_g5 and c.has_flatsky_nonlocal
    c.to_file(tmp_path / "back.cfg")
    d = Coffe.from_file(tmp_path / "back.cfg")
    assert d.has_lensing and d.has_g4 and d.has_g5 and not d.has_rsd


def test_bias_file_is_read(tmp_path):
    zb = np.linspace(0, 4, 9)
    np.savetxt(tmp_path / "b.dat", np.transpose([zb, 2 - 0.1 * zb]))
    cfg = tmp_path / "b.cfg"
    cfg.write_text(f'read_galaxy_bias2 = 1;\ninput_galaxy_bias2 = "{tmp_path / "b.dat"}";\n')
    c = Coffe.from_file(cfg)
    assert np.allclose(c.galaxy_bias2_redshifts(), zb) and np.allclose(c.galaxy_bias2_values(), 2 - 0.1 * zb)


def test_bias_samples_are_even():
    """The getters give the spline's size evenly spaced points and its values there."""
    c = Coffe()
    x = np.array([0.0, 0.1, 0.5, 1.0, 2.0, 4.0])
    c.set_evolution_bias1(x, x ** 2)
    zs = c.evolution_bias1_redshifts()
    assert np.allclose(zs, np.linspace(0, 4, 6))
    vals = c.evolution_bias1_values()
    assert vals[0] == pytest.approx(0.0) and vals[-1] == pytest.approx(16.0)
    assert np.allclose(vals, [c.evolution_bias1(float(z)) for z in zs])


def test_choice_parameters():
    c = Coffe()
    c.covariance_integration_method = 2
    assert c.covariance_integration_method == "fftlog"
    c.covariance_interpolation_method = "bilinear"
    assert c.covariance_interpolation_method == "bilinear"
    c.integration_1d_type = "exponential"
    assert c.integration_1d_type == "exponential"
    with pytest.raises(ValueError):
        c.integration_1d_type = "simpson"
    c.pk_type = 0
    assert c.pk_type == "linear"
    with pytest.raises(NotImplementedError):
        c.pk_type = "halofit"
    with pytest.raises(NotImplementedError):
        c.m_ncdm = 0.06
    c.has_flatsky_nonlocal = True
    c.has_flatsky_local_nonlocal = False
    assert c.has_flatsky_nonlocal and not c.has_flatsky_local_nonlocal


def test_kept_parameters_are_checked():
    c = Coffe()
    c.sigma8 = 0.9
    assert c.sigma8 == 0.9 and c.parameters["sigma8"] == 0.9
    with pytest.raises(ValueError):
        c.n_s = 101
    with pytest.raises(TypeError):
        c.N_ncdm = 1.5
    with pytest.raises(TypeError):
        c.integration_1d_prec = 1          # a float in the reference
    with pytest.raises(ValueError):
        c.k_min = 0.1
    with pytest.raises(ValueError):
        c.k_max = 0.01
    with pytest.raises(TypeError):
        c.background_sampling = 100.0


def test_k_window_follows_the_spectrum():
    c = Coffe()
    k = np.logspace(-4, 1, 50)
    c.set_power_spectrum_linear(k, k ** -1)
    assert c.k_min == k[0] and c.k_max == k[-1]
    assert c._kwin() == (k[0], k[-1])
    c.k_min = 5e-4
    assert c._kwin() == (5e-4, k[-1])
    c.k_min = 1e-5                         # below the samples
    with pytest.raises(ValueError):
        c._kwin()


def test_omega_bookkeeping():
    """omega_baryon and omega_m keep the other of the pair and move omega_cdm
    (coffe.pyx:1207-1236)."""
    c = Coffe()
    c.omega_baryon = 0.04
    assert c.omega_m == pytest.approx(0.3) and c.omega_cdm == pytest.approx(0.26)
    c.omega_m = 0.32
    assert c.omega_baryon == 0.04 and c.omega_cdm == pytest.approx(0.28)
    c.omega_cdm = 0.2
    assert c.omega_m == pytest.approx(0.24)
    assert c.big_omega_m == c.omega_m and c.big_omega_de == pytest.approx(1 - 0.24 - 9e-5)
    assert c.omega_nu == 0.0 and not c.use_little_omega and not c.has_class
    with pytest.raises(ValueError):
        c.omega_gamma = 0.8


def test_bilinear_covariance_is_refused():
    c = Coffe(covariance_interpolation_method="bilinear")
    k = np.logspace(-4, 1, 50)
    c.set_power_spectrum_linear(k, k ** -1)
    with pytest.raises(NotImplementedError):
        c.compute_covariance_bulk()
