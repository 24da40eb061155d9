"""Derivatives of the multipoles with respect to the background parameters through the
engine's forward-mode weights (coffe_mu_weights_jvp, Coffe.multipoles_cosmo_jvp): full
and flat sky, density + RSD + Doppler-1 with their cross terms, against central
differences of the whole Coffe pipeline (background, weights, projection) at
theta +/- eps d, relative step 1e-4 (truncation ~1e-8): 1e-6 of the largest |dM| at
each redshift.  The weights of the dual kernel equal the plain kernel's (1e-13)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
BASE = dict(omega_cdm=0.25, omega_baryon=0.05, omega_gamma=9e-5, w0=-1.0, wa=0.0, h=0.67)


def _pk():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k * H, pk / H ** 3


def _cosmo(flat, **theta):
    """A Coffe object at theta = (omega_cdm, omega_baryon, ...).  Its omega_baryon setter
    holds omega_m (coffe.pyx:1229-1236), so omega_baryon goes first and omega_m = cdm + b
    after it, which leaves omega_cdm at theta's value."""
    from coffe_amd.coffe import Coffe
    rest = dict(theta)
    cdm, baryon = rest.pop("omega_cdm"), rest.pop("omega_baryon")
    theta = dict(omega_baryon=baryon, omega_m=cdm + baryon, **rest)
    c = Coffe(sep=np.array([15.0, 45.0, 110.0, 220.0]), l=[0, 1, 2, 3, 4], z_mean=[0.6, 1.4], has_density=True,
              has_rsd=True, has_d1=True, has_flatsky_local=flat, integrals_fftlog_sampling=4096, **theta)
    assert c.omega_cdm == pytest.approx(cdm, rel=1e-15, abs=1e-15)
    c.set_power_spectrum_linear(*_pk())
    zb = np.linspace(0, 15, 16)
    c.set_galaxy_bias1(zb, 1.1 + 0.4 * zb)
    c.set_galaxy_bias2(zb, 0.8 + 0.6 * zb)
    return c


@pytest.mark.parametrize("flat", [False, True])
@pytest.mark.parametrize("direction", [{"omega_cdm": 1.0}, {"w0": 1.0}, {"h": 1.0},
                                       {"omega_baryon": 1.0, "wa": 0.5}])
def test_cosmo_jvp_vs_central_differences(engine, flat, direction):
    c = _cosmo(flat, **BASE)
    print(f"multipoles_cosmo_jvp flat={flat} {direction}", flush=True)
    jv = c.multipoles_cosmo_jvp(direction).reshape(2, 5, 4)
    eps = 1e-4
    vals = []
    for sgn in (1, -1):
        theta = {k: v + sgn * eps * direction.get(k, 0.0) * max(abs(v), 1.0) for k, v in BASE.items()}
        vals.append(np.array([r.value for r in _cosmo(flat, **theta).compute_multipoles_bulk()]).reshape(2, 5, 4))
    # the step along parameter p was eps max(|p|, 1) direction[p]: rescale to d/d theta
    scale = {k: max(abs(v), 1.0) for k, v in BASE.items()}
    steps = {scale[k] for k in direction}
    assert len(steps) == 1, "directions mixing parameters of different step scales"
    fd = (vals[0] - vals[1]) / (2 * eps * steps.pop())
    # per redshift, against the largest derivative of any multipole there: a multipole
    # that vanishes analytically (the flat-sky octupole with G1 = G2) is round-off in both
    for iz in range(2):
        ref = np.max(np.abs(fd[iz]))
        for il in range(5):
            assert np.max(np.abs(jv[iz, il] - fd[iz, il])) <= 1e-6 * ref, (flat, direction, iz, il,
                                                                              jv[iz, il], fd[iz, il])


def test_dual_weights_equal_plain(engine):
    from coffe_amd import background as B
    from coffe_amd.coffe import TABLES
    from coffe_amd.engine import mu_weights, mu_weights_jvp
    bg, tg = B.jvp({"omega_cdm": 1.0})
    tabs = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"], G2=bg["G2"])
    tans = dict(chi=tg["chi"], D1=tg["D1"], f=tg["f"], cH=tg["conformal_Hz"], G1=tg["G1"], G2=tg["G2"])
    zb = np.linspace(0, 10, 11)
    b1, b2 = 1.2 + 0.1 * zb, 1.5 + 0.2 * zb
    sep = np.array([10.0, 50.0, 200.0])
    for flat in (False, True):
        A = mu_weights(engine, tabs, zb, b1, b2, [0.5, 1.0], [0, 2, 4], list(TABLES), sep, den=True, rsd=True, d1=True,
                       flatsky=flat).cpu().numpy()
        Ad, At = mu_weights_jvp(engine, tabs, tans, zb, b1, b2, [0.5, 1.0], [0, 2, 4], list(TABLES), sep, den=True,
                                rsd=True, d1=True, flatsky=flat)
        Ad = Ad.cpu().numpy()
        assert np.max(np.abs(Ad - A)) <= 1e-13 * np.max(np.abs(A)), flat
        assert np.all(np.isfinite(At.cpu().numpy()))
