"""Segmented detectors and geometric-correction tables in the host geometry (xpg/geometry.py:185-304,
xpg/detector.py:1-170), pinned without a GPU: one full segment equals the plain panel; raw pixel -> ray -> raw pixel
round trips on two tilted, shifted segments with distortion tables and parallax; gap pixels belong to no segment; the
oracle's kernel-side lab position (oracle.kernels.pixel_S, the arithmetic of the HIP pixel_S) agrees with the host
geometry, as the reference's own test requires of its kernels (tests/test_features.py:62-73 of the reference: 2068 x
2162 pixels, a 0.4 degree tilt, 0.45 mm sensor with mu 3.7, relative error below 1e-6)."""
import numpy as np
import pytest

from oracle import kernels as ok
from xpg_amd.geometry import Geometry, Segment, detector_consts


def _geom(**kw):
    kw = dict(dict(sensor_thickness=0.45, mu=3.7), **kw)
    return Geometry(wavelength=1.0, beam=np.array([0, 0, 1.0]), axis=np.array([1.0, 0, 0]), orgx=1030, orgy=1080,
                    distance=150, ed1=np.array([1.0, 0, 0]), ed2=np.array([0, 1.0, 0]), qx=0.075, qy=0.075, nx=2068,
                    ny=2162, phi0=0, osc=0.2, starting_frame=1, **kw)


def _two_segments(tilt_z=0.02):
    th = np.radians(0.4)
    return [Segment(1, 2068, 1, 1000),
            Segment(1, 2068, 1040, 2162, orgx=0.4, orgy=-0.3, distance=0.2,
                    ex=np.array([np.cos(th), np.sin(th), tilt_z]), ey=np.array([-np.sin(th), np.cos(th), 0.0]))]


def _distortion():
    yy, xx = np.mgrid[0:2162, 0:2068]
    return (0.5 * np.sin(xx / 150.0)).astype(np.float32), (0.4 * np.cos(yy / 200.0)).astype(np.float32)


def test_one_full_segment_is_the_plain_panel():
    g = _geom()
    g1 = _geom(segments=[Segment(1, 2068, 1, 2162)])
    rng = np.random.default_rng(0)
    x, y = rng.uniform(1, 2068, 500), rng.uniform(1, 2162, 500)
    assert np.abs(g1.pixel_to_lab(x, y) - g.pixel_to_lab(x, y)).max() < 1e-9
    P0, Bx, By = g.panels()
    np.testing.assert_allclose(P0, [[0, 0, 150.0]], atol=1e-12)
    np.testing.assert_allclose(Bx, [[0.075, 0, 0]], atol=1e-15)
    np.testing.assert_allclose(By, [[0, 0.075, 0]], atol=1e-15)


def test_segment_map_and_gaps():
    g = _geom(segments=_two_segments())
    m = g.segmap()
    assert m.shape == (2162, 2068) and m.dtype == np.int16
    assert (m[:1000] == 0).all() and (m[1000:1039] == -1).all() and (m[1039:] == 1).all()
    assert g.segment_of(np.array([10.0, 10.0, 10.0]), np.array([999.9, 1020.0, 1040.2])).tolist() == [0, -1, 1]


@pytest.mark.parametrize("distortion", [False, True])
def test_pixel_ray_pixel_round_trip(distortion):
    g = _geom(segments=_two_segments(), geo_corr=_distortion() if distortion else None)
    rng = np.random.default_rng(1)
    x, y = rng.uniform(5, 2063, 2000), rng.uniform(5, 2157, 2000)
    k = g.segment_of(x, y)
    m = k >= 0
    assert (~m).sum() > 0
    xb, yb = g.s_to_pixel(g.pixel_to_s(x[m], y[m]))
    assert np.abs(xb - x[m]).max() < 1e-6 and np.abs(yb - y[m]).max() < 1e-6


def test_second_segment_moves_the_lab_position():
    # the shifted segment's pixel is offset by its SEGMENT_ORGX / ORGY (pixels) and SEGMENT_DISTANCE (mm)
    seg = Segment(1, 2068, 1, 2162, orgx=2.0, orgy=-1.0, distance=0.5)
    g, g1 = _geom(sensor_thickness=0.0), _geom(segments=[seg], sensor_thickness=0.0)
    X, X1 = g.pixel_to_lab(500.0, 600.0), g1.pixel_to_lab(500.0, 600.0)
    np.testing.assert_allclose(X1 - X, [-2.0 * 0.075, 1.0 * 0.075, 0.5], atol=1e-12)


@pytest.mark.parametrize("segmented", [False, True])
def test_kernel_lab_position_matches_host_geometry(segmented):
    g = _geom(segments=_two_segments(tilt_z=0.0) if segmented else None)
    d = detector_consts(g)
    assert (d["segmap"] is not None) == segmented and d["P0"].shape == ((2, 3) if segmented else (1, 3))
    rng = np.random.default_rng(2)
    x, y = rng.uniform(1, 2068, 300), rng.uniform(1, 2162, 300)
    Sk = ok.pixel_S(d, x, y)
    Sh = g.pixel_to_s(x, y)
    assert np.abs(Sk - Sh).max() / np.abs(Sh).max() < 1e-6


def test_kernel_lab_position_with_distortion_tables():
    g = _geom(segments=_two_segments(tilt_z=0.0), geo_corr=_distortion(), sensor_thickness=0.0)
    d = detector_consts(g)
    assert d["gx"].dtype == np.float32 and d["gx"].shape == (2162, 2068)
    rng = np.random.default_rng(3)
    x, y = rng.uniform(1, 2068, 300), rng.uniform(1, 2162, 300)
    # without parallax the kernel arithmetic and the host geometry are the same map
    assert np.abs(ok.pixel_S(d, x, y) - g.pixel_to_s(x, y)).max() / np.abs(g.pixel_to_s(x, y)).max() < 1e-12
