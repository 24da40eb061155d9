"""The command line on the GPU: its tables hold what the Coffe API computes for the
same settings (python/coffe/__main__.py:86-160)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67


@pytest.fixture
def setup(tmp_path):
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    pkfile = tmp_path / "pk.dat"
    np.savetxt(pkfile, np.transpose([k * H, pk / H ** 3]))
    cfg = tmp_path / "s.cfg"
    cfg.write_text("has_density = 1;\nhas_rsd = 1;\nsep = [20.0, 40.0, 80.0];\nl = [0, 2];\nz_mean = [1.0];\n"
                   "deltaz = [0.2];\nz_min = [0.9];\nz_max = [1.1];\nmu = [0.3, 0.8];\n"
                   "number_density1 = [1e-3];\nnumber_density2 = [1e-3];\nfsky = [0.5];\npixelsize = [5.0];\n")
    from coffe_amd.coffe import Coffe
    cosmo = Coffe.from_file(cfg)
    cosmo.set_power_spectrum_linear(k * H, pk / H ** 3)
    return cfg, pkfile, cosmo


@pytest.mark.parametrize("kind", ["mp", "cf", "cov", "avg_mp", "avg_cov"])
def test_cli_tables_match_api(setup, tmp_path, kind):
    from coffe_amd.__main__ import main
    cfg, pkfile, cosmo = setup
    out = tmp_path / f"{kind}.dat"
    print(f"cli kind={kind}", flush=True)
    main(["-s", str(cfg), "-o", str(out), "-k", kind, "-p", str(pkfile)])
    data = np.atleast_2d(np.loadtxt(out, delimiter="\t"))
    if kind == "mp":
        ref = [[r.r, r.l, r.z, r.value] for r in cosmo.compute_multipoles_bulk()]
    elif kind == "cf":
        ref = [[r.r, r.mu, r.z, r.value] for r in cosmo.compute_corrfunc_bulk()]
    elif kind == "cov":
        ref = [[r.r1, r.r2, r.l1, r.l2, r.z, r.value] for r in cosmo.compute_covariance_bulk()]
    elif kind == "avg_mp":
        ref = [[r.r, r.l, r.z_min, r.z_max, r.value] for r in cosmo.compute_average_multipoles_bulk()]
    else:
        ref = [[r.r1, r.r2, r.l1, r.l2, r.z_min, r.z_max, r.value] for r in cosmo.compute_average_covariance_bulk()]
    ref = np.array(ref)
    assert data.shape == ref.shape
    # "%.10e": ten digits after the point
    assert np.allclose(data, ref, rtol=1e-10, atol=0)
