"""The text writers against the formats of output.c / common.c and of the Python
command line (python/coffe/__main__.py), on the CPU."""
import os
import re

import numpy as np
import pytest

from coffe_amd import output
from coffe_amd.coffe import Coffe
from coffe_amd.representation import AverageCovariance, Corrfunc, Covariance, Multipoles

GOLD = os.path.join(os.path.dirname(__file__), "golden")
E10 = r"-?\d\.\d{10}e[+-]\d{2}"


def test_multipoles_file(tmp_path):
    recs = [Multipoles(z=1.0, l=0, r=5.0, value=1.25e-2), Multipoles(z=1.0, l=2, r=8.0, value=-3.5e-4)]
    path = tmp_path / "multipoles.dat"
    output.write_multipoles(path, recs, dict(has_density=True, has_rsd=True, has_g2=True))
    assert path.read_text() == ("# correlation contributions: den rsd g2 \n# z_mean, l, sep, value\n"
                                "1.0000000000e+00 0 5.0000000000e+00 1.2500000000e-02\n"
                                "1.0000000000e+00 2 8.0000000000e+00 -3.5000000000e-04\n")
    # the C writer appends (output.c:152)
    output.write_multipoles(path, recs[:1], dict(has_d1=True))
    assert path.read_text().count("# z_mean, l, sep, value") == 2 and "contributions: d1 \n" in path.read_text()


def test_corrfunc_and_covariance_files(tmp_path):
    output.write_corrfunc(tmp_path / "cf.dat", [Corrfunc(z=0.5, r=10.0, mu=0.7, value=2e-3)], dict(has_lensing=True))
    assert (tmp_path / "cf.dat").read_text() == ("# correlation contributions: len \n# z_mean, sep, mu, value\n"
                                                 "5.0000000000e-01 1.0000000000e+01 7.0000000000e-01 "
                                                 "2.0000000000e-03\n")
    output.write_covariance(tmp_path / "cov.dat", [Covariance(z=1.0, l1=0, l2=2, r1=50.0, r2=60.0, value=8.1e-8)],
                            dict(has_density=True))
    assert (tmp_path / "cov.dat").read_text().splitlines() == [
        "# correlation contributions: den ", "# z_mean, l1, l2, sep1, sep2, value",
        "1.0000000000e+00 0 2 5.0000000000e+01 6.0000000000e+01 8.1000000000e-08"]
    output.write_average_covariance(
        tmp_path / "avg.dat", [AverageCovariance(r1=50.0, r2=60.0, l1=2, l2=4, z_min=0.9, z_max=1.1, value=1e-9)])
    assert (tmp_path / "avg.dat").read_text() == ("# sep[Mpc/h]\tsep[Mpc/h]\tresult\n"
                                                  "5.0000000000e+01 6.0000000000e+01 2 4 9.0000000000e-01 "
                                                  "1.1000000000e+00 1.0000000000e-09\n")


def test_background_file(tmp_path):
    """write_ncol_null (common.c:398-405) ends every value with the separator, the last
    one included; the header names the columns with their units (output.c:201-300)."""
    c = Coffe(background_sampling=500)
    bg = c._background()
    path = tmp_path / "background.dat"
    output.write_background(path, bg)
    lines = path.read_text().split("\n")
    assert lines[0] == ("# 1)z\t2)a\t3)H[1/Mpc]\t4)conformal_H[1/Mpc]\t5)conformal_H_prime[1/Mpc^2]\t6)D1\t7)f\t"
                        "8)comoving_distance[Mpc]\t")
    assert len(lines) == 502 and lines[-1] == ""
    assert re.fullmatch("(" + E10 + "\t){8}", lines[1])
    data = np.loadtxt(path)
    assert data.shape == (500, 8)
    for col, key in zip(data.T, ("z", "a", "Hz", "conformal_Hz", "conformal_Hz_prime", "D1", "f", "chi")):
        assert np.allclose(col, bg[key], rtol=1e-10, atol=0)
    assert data[0, 0] == 0.0 and data[-1, 0] == 15.0
    with pytest.raises(ValueError):
        output.write_background(path, bg, columns=("z", "G1"))


def test_integral_files(tmp_path):
    x = np.array([1.0, 2.0, 3.0])
    tables = {nl: (x, x * (i + 1)) for i, nl in enumerate(output.INTEGRAL_FILES)}
    chi = np.array([0.0, 10.0])
    grid = np.array([[1.0, 2.0], [3.0, 4.0]])
    output.write_integrals(str(tmp_path) + "/run_", tables, divergent=(x, -x), renormalization=(chi, grid))
    assert (tmp_path / "run_integral3.dat").read_text() == ("# n = 1, l = 1\n1.0000000000e+00 4.0000000000e+00 \n"
                                                            "2.0000000000e+00 8.0000000000e+00 \n"
                                                            "3.0000000000e+00 1.2000000000e+01 \n")
    assert (tmp_path / "run_integral8.dat").read_text().startswith("# n = 4, l = 0\n1.0000000000e+00 "
                                                                    "-1.0000000000e+00 \n")
    # grid[j, i] is the value at (chi_i, chi_j); "%e" per value (output.c:527-545)
    assert (tmp_path / "run_integral8_renormalization.dat").read_text() == (
        "0.000000e+00 0.000000e+00 1.000000e+00\n0.000000e+00 1.000000e+01 3.000000e+00\n"
        "1.000000e+01 0.000000e+00 2.000000e+00\n1.000000e+01 1.000000e+01 4.000000e+00\n")


def test_output_prefix(tmp_path):
    assert output.output_prefix(str(tmp_path / "a" / "b"), "run_") == str(tmp_path / "a" / "b") + "/run_"
    assert os.path.isdir(tmp_path / "a" / "b")
    assert output.output_prefix(str(tmp_path) + "/", "$TIME", timestamp="T0") == str(tmp_path) + "/T0_"


def test_cli_background(tmp_path):
    """python -m coffe_amd -k bg: the background at background_sampling redshifts in
    [0, 15], tab separated under one header line (python/coffe/__main__.py:55-85)."""
    from coffe_amd.__main__ import main
    cfg = tmp_path / "s.cfg"
    cfg.write_text("background_sampling = 300\nomega_m = 0.31\n")
    out = tmp_path / "bg.dat"
    main(["-s", str(cfg), "-o", str(out), "-k", "bg"])
    lines = out.read_text().splitlines()
    assert lines[0] == ("# 1)z\t2)a\t3)H[1/Mpc]\t4)conformal_H[1/Mpc]\t5)conformal_H_prime[1/Mpc^2]\t6)D1\t7)f\t"
                        "8)comoving_distance[Mpc]")
    data = np.loadtxt(out, delimiter="\t")
    assert data.shape == (300, 8)
    c = Coffe(omega_m=0.31, background_sampling=300)
    assert np.allclose(data[:, 7], c._background()["chi"], rtol=1e-10)
    assert data[57, 6] == pytest.approx(c.growth_rate(float(data[57, 0])), rel=1e-10)


def test_cli_needs_power_spectrum(tmp_path):
    from coffe_amd.__main__ import main
    cfg = tmp_path / "s.cfg"
    cfg.write_text("l = [0]\n")
    with pytest.raises(SystemExit):
        main(["-s", str(cfg), "-o", str(tmp_path / "mp.dat"), "-k", "mp"])
