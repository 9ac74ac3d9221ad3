"""Pins of the seeded balanced-jet initial state (mcv_host_jet in the Fortran
host library), the generator every configuration starts from.

The reference has no such generator; its zero-perturbation limit is the flow
of the isolated-mountain case without the mountain (test_case_mod.F90:
116-120).  The checks here are fixed by the mathematics, not by the code:
  - zero perturbation: the closed-form balanced jet, and Williamson case 5's
    phi + phi_s (the same flow over its mountain);
  - the perturbation is linear in its amplitude and has zero sphere mean;
  - its sphere-mean rms equals rms_h, by Gauss-Legendre x trapezoid
    quadrature that is exact for the band-limited h'^2 (this holds only if
    the Y_lm are orthonormal, for every degree up to lmax);
  - degree 1 alone is a linear function of the Cartesian coordinates;
  - the perturbation wind is non-divergent (a streamfunction wind);
  - the same seed gives the same state, a different seed another one.
The generator takes arbitrary (lon, lat) points, so the quadrature and the
finite differences are evaluated on grids of their own.
"""
import numpy as np
import pytest

from mcv_sw_amd import hostlib

G, R, OM = hostlib.GRAVITY, hostlib.RADIUS, hostlib.OMEGA


def _jet_at(lon, lat, **kw):
    """mcv_host_jet on arbitrary points; returns (phi, zonal, merid)."""
    lon = np.ascontiguousarray(lon, dtype=np.float64).ravel()
    lat = np.ascontiguousarray(lat, dtype=np.float64).ravel()
    m = {"lonP": lon, "latP": lat}
    phi, z, w, hs = (np.empty(lon.size) for _ in range(4))
    args = dict(u0=20.0, h0=5960.0, lmax=8, seed=0, rms_h=30.0)
    args.update(kw)
    rc = hostlib.load().mcv_host_jet(lon.size, R, OM, G, args["u0"], args["h0"], args["lmax"],
                                     args["seed"], args["rms_h"], hostlib._p(lon),
                                     hostlib._p(lat), hostlib._p(phi), hostlib._p(z),
                                     hostlib._p(w), hostlib._p(hs))
    assert rc == 0
    assert not hs.any()
    del m
    return phi, z, w


def _gauss_grid(lmax):
    """Nodes and weights exact for products of two degree-lmax fields:
    Gauss-Legendre in sin(lat), uniform in longitude."""
    nlat = lmax + 2
    x, wx = np.polynomial.legendre.leggauss(nlat)
    nlon = 2 * lmax + 3
    lon = (np.arange(nlon) + 0.25) * (2.0 * np.pi / nlon)
    LON, LAT = np.meshgrid(lon, np.arcsin(x))
    W = np.repeat(wx[:, None], nlon, axis=1) * (2.0 * np.pi / nlon)  # sums to 4 pi
    return LON, LAT, W


def _perturbation(lon, lat, **kw):
    phi, z, w = _jet_at(lon, lat, **kw)
    phi0, z0, w0 = _jet_at(lon, lat, **{**kw, "rms_h": 0.0})
    return (phi - phi0) / G, z - z0, w - w0


def test_zero_perturbation_is_the_balanced_jet():
    rng = np.random.default_rng(1)
    lon = rng.uniform(-np.pi, np.pi, 4000)
    lat = np.arcsin(rng.uniform(-1.0, 1.0, 4000))
    u0, h0 = 40.0, 8000.0
    phi, z, w = _jet_at(lon, lat, u0=u0, h0=h0, rms_h=0.0, lmax=16, seed=7)
    phi_cf = G * h0 - (R * OM * u0 + u0 ** 2 / 2.0) * np.sin(lat) ** 2
    assert np.abs(phi - phi_cf).max() <= 4 * np.finfo(float).eps * G * h0
    assert np.abs(z - u0 * np.cos(lat)).max() <= 4 * np.finfo(float).eps * u0
    assert not w.any()


def test_zero_perturbation_is_case5_without_its_mountain():
    """test_case_mod.F90:116-120: gh = g h0 - (R Omega u0 + u0^2/2) sin^2(lat) - hs,
    u = u0 cos(lat), v = 0 (alpha = 0); the jet with u0 = 20, h0 = 5960 is that
    flow with hs = 0, i.e. case 5's phi + phi_s."""
    m = hostlib.mesh(3, 8)
    c5 = hostlib.williamson(m, 5)
    hs = m["phi_s"].copy()
    m2 = dict(m)
    jet = hostlib.jet(m2, u0=20.0, h0=5960.0, lmax=8, seed=0, rms_h=0.0)
    assert np.abs(jet["phi"] - (c5["phi"] + hs)).max() <= 4 * np.finfo(float).eps * G * 5960.0
    assert np.array_equal(jet["zonal"], c5["zonal"])
    assert not jet["merid"].any()


@pytest.mark.parametrize("lmax,rms", [(1, 30.0), (8, 30.0), (12, 20.0), (16, 60.0)])
def test_perturbation_rms_is_rms_h(lmax, rms):
    LON, LAT, W = _gauss_grid(lmax)
    for seed in (0, 1, 5):
        hp, _, _ = _perturbation(LON, LAT, lmax=lmax, seed=seed, rms_h=rms)
        hp = hp.reshape(LON.shape)
        mean = (W * hp).sum() / (4.0 * np.pi)
        ms = (W * hp * hp).sum() / (4.0 * np.pi)
        assert abs(mean) <= 1e-11 * rms  # no degree-0 part
        assert abs(np.sqrt(ms) - rms) <= 1e-11 * rms


def test_perturbation_is_linear_in_its_amplitude():
    rng = np.random.default_rng(2)
    lon = rng.uniform(-np.pi, np.pi, 500)
    lat = np.arcsin(rng.uniform(-1.0, 1.0, 500))
    a = _perturbation(lon, lat, lmax=12, seed=3, rms_h=10.0)
    b = _perturbation(lon, lat, lmax=12, seed=3, rms_h=40.0)
    for x, y in zip(a, b):
        assert np.abs(y - 4.0 * x).max() <= 1e-9 * np.abs(y).max()


def test_degree_one_is_linear_in_cartesian_coordinates():
    """Degree-1 real harmonics span {x, y, z}: h' = c . (x, y, z) exactly."""
    rng = np.random.default_rng(3)
    lon = rng.uniform(-np.pi, np.pi, 300)
    lat = np.arcsin(rng.uniform(-1.0, 1.0, 300))
    hp, _, _ = _perturbation(lon, lat, lmax=1, seed=11, rms_h=25.0)
    X = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=1)
    c, *_ = np.linalg.lstsq(X, hp, rcond=None)
    assert np.abs(X @ c - hp).max() <= 1e-12 * np.abs(hp).max()
    # the rms of c . r over the sphere is |c| / sqrt(3)
    assert abs(np.linalg.norm(c) / np.sqrt(3.0) - 25.0) <= 1e-11 * 25.0


def test_perturbation_wind_is_non_divergent():
    """div v = 1/(a cos) (du/dlon + d(v cos)/dlat) vanishes for a
    streamfunction wind; centred differences of the generator's own winds
    give a divergence at the truncation level, far below the vorticity."""
    lmax, rms = 8, 30.0
    lat0 = np.linspace(-1.3, 1.3, 27)
    lon0 = np.linspace(-3.0, 3.0, 31)
    LON, LAT = np.meshgrid(lon0, lat0)
    e = 1e-4

    def wind(lon, lat):
        _, u, v = _perturbation(lon, lat, lmax=lmax, seed=2, rms_h=rms)
        return u, v

    up, _ = wind(LON + e, LAT)
    um, _ = wind(LON - e, LAT)
    _, vp = wind(LON, LAT + e)
    _, vm = wind(LON, LAT - e)
    div = ((up - um) / (2 * e) + (vp * np.cos(LAT.ravel() + e) - vm * np.cos(LAT.ravel() - e))
           / (2 * e)) / (R * np.cos(LAT.ravel()))
    vp2, _ = wind(LON, LAT + e)
    vm2, _ = wind(LON, LAT - e)
    _, wp = wind(LON + e, LAT)
    _, wm = wind(LON - e, LAT)
    vort = ((wp - wm) / (2 * e) - (vp2 * np.cos(LAT.ravel() + e) - vm2 * np.cos(LAT.ravel() - e))
            / (2 * e)) / (R * np.cos(LAT.ravel()))
    assert np.abs(vort).max() > 0
    assert np.abs(div).max() <= 1e-6 * np.abs(vort).max()


def test_seed_reproducibility():
    m = hostlib.mesh(3, 6)
    a = hostlib.jet(dict(m), u0=20.0, h0=5960.0, lmax=12, seed=42, rms_h=20.0)
    b = hostlib.jet(dict(m), u0=20.0, h0=5960.0, lmax=12, seed=42, rms_h=20.0)
    c = hostlib.jet(dict(m), u0=20.0, h0=5960.0, lmax=12, seed=43, rms_h=20.0)
    for k in ("phi", "zonal", "merid", "u", "v", "phiG", "contraU", "contraV"):
        assert np.array_equal(a[k], b[k]), k
    assert not np.array_equal(a["phi"], c["phi"])
    # members of one ensemble draw different perturbations of equal rms
    LON, LAT, W = _gauss_grid(12)
    r = [np.sqrt((W * _perturbation(LON, LAT, lmax=12, seed=s, rms_h=20.0)[0].reshape(LON.shape)
                  ** 2).sum() / (4 * np.pi)) for s in (1, 2, 64)]
    assert np.allclose(r, 20.0, rtol=1e-11, atol=0)
