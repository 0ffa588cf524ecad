"""Host-side pins of the product (no GPU): the aperture generators of hcipy_amd.aperture against
the reference's golden rasters (tests/test_aperture.py:8-40, tests/baseline_for_apertures/**),
and the FIFO instance cache of AgnosticOpticalElement
(hcipy/optics/optical_element.py:290-380)."""
import glob
import os

import numpy as np
import pytest

import hcipy_amd as ha
from hcipy_amd.optics import AgnosticOpticalElement, wavelength_key
from oracle import hcipy_oracle as o

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'apertures')


def _host(f):
    return np.asarray(f.as_subclass(__import__('torch').Tensor).cpu()) if hasattr(f, 'as_subclass') else np.asarray(f)


@pytest.mark.parametrize('path', sorted(glob.glob(os.path.join(GOLDEN, '*.fits.gz'))))
def test_product_apertures_match_reference_rasters(path):
    """As check_aperture (tests/test_aperture.py:8-40): 256^2 over the unit square, supersampled
    8x, np.allclose against the stored raster."""
    name = os.path.basename(path)
    grid = ha.make_uniform_grid(256, [1, 1])
    if name.startswith('pupil_'):
        ap = ha.make_circular_aperture(0.5 if 'small' in name else 1.0)
    else:
        ratio = 0.1 if 'small_obscuration' in name else 0.3
        spiders = 3 if '3spiders' in name else 5
        width = 0.01 if 'thin' in name else 0.03
        ap = ha.make_obstructed_circular_aperture(1, ratio, spiders, width)
    field = _host(ha.evaluate_supersampled(ap, grid, 8))
    ref = o.read_fits_image(path).ravel()
    assert field.shape == ref.shape
    assert np.allclose(field, ref)


@pytest.mark.parametrize('path', sorted(glob.glob(os.path.join(GOLDEN, 'super_gaussian', '*.fits.gz'))))
def test_product_super_gaussian_matches_reference_rasters(path):
    """tests/test_aperture.py:99-114 on the product generator."""
    name = os.path.basename(path)
    fwhm = 0.5 if '_small_' in name else 1.0
    p = 1.0 if 'smallshape' in name else 10.0
    center = (0.25, -0.3) if 'shifted' in name else None
    grid = ha.make_uniform_grid(256, [1, 1])
    field = _host(ha.evaluate_supersampled(ha.make_super_gaussian_aperture(fwhm, p, center), grid, 8))
    assert np.allclose(field, o.read_fits_image(path).ravel())


def test_spider_alone_against_a_rotated_box():
    """make_spider: a bar of the given width between p1 and p2 is 0, everything else 1; checked
    against the box written in the bar's own frame for a diagonal spider."""
    grid = ha.make_uniform_grid(64, [1, 1])
    p1, p2, w = (-0.3, -0.2), (0.35, 0.25), 0.08
    f = _host(ha.make_spider(p1, p2, w)(grid))
    x, y = grid.x, grid.y
    d = np.subtract(p2, p1)
    t = np.arctan2(d[1], d[0])
    cx, cy = (p1[0] + p2[0]) / 2, (p1[1] + p2[1]) / 2
    u = (x - cx) * np.cos(t) + (y - cy) * np.sin(t)
    v = -(x - cx) * np.sin(t) + (y - cy) * np.cos(t)
    inside = (np.abs(u) <= np.hypot(*d) / 2) & (np.abs(v) <= w / 2)
    assert inside.sum() > 20
    assert np.array_equal(f, np.where(inside, 0.0, 1.0))


class _Counting(AgnosticOpticalElement):
    def __init__(self, **kw):
        AgnosticOpticalElement.__init__(self, **kw)
        self.made = []

    def make_instance(self, data, input_grid, output_grid, wavelength):
        self.made.append(wavelength)
        data.tag = len(self.made)


def test_instance_cache_is_fifo_with_max_in_cache():
    """The oldest instance leaves first once max_in_cache are held, and a hit does not refresh an
    instance's place (optical_element.py:358-380: OrderedDict.popitem(last=False) on insertion
    order)."""
    grid = ha.make_pupil_grid(8)
    el = _Counting(max_in_cache=3)
    for wl in (1.0, 2.0, 3.0):
        el.get_instance_data(grid, grid, wl)
    assert el.made == [1.0, 2.0, 3.0]
    el.get_instance_data(grid, grid, 1.0)  # hit: no new instance, no refresh
    assert el.made == [1.0, 2.0, 3.0]
    el.get_instance_data(grid, grid, 4.0)  # evicts 1.0 (oldest inserted), not 2.0
    assert el.made == [1.0, 2.0, 3.0, 4.0]
    el.get_instance_data(grid, grid, 2.0)
    assert el.made == [1.0, 2.0, 3.0, 4.0]
    el.get_instance_data(grid, grid, 1.0)  # rebuilt, evicting 2.0
    assert el.made[-1] == 1.0 and len(el.made) == 5
    el.get_instance_data(grid, grid, 2.0)
    assert len(el.made) == 6
    el.clear_cache()
    el.get_instance_data(grid, grid, 4.0)
    assert len(el.made) == 7


def test_instance_cache_keys():
    """Wavelengths share an instance when round(log(wl) / log(1 + 1e-9)) agrees (optical_element.py:333);
    grids are keyed by content, so an equal grid built again hits; grid- and wavelength-independent
    elements ignore those parts of the key; the default holds 64 instances."""
    g1 = ha.make_pupil_grid(16)
    g2 = ha.make_pupil_grid(16)
    g3 = ha.make_pupil_grid(16, 2.0)
    el = _Counting()
    assert el.max_in_cache == 64
    wl = 5e-7
    a = el.get_instance_data(g1, g1, wl)
    assert el.get_instance_data(g2, g2, wl * (1 + 1e-12)) is a
    assert wavelength_key(wl) == int(np.round(np.log(wl) / np.log(1 + 1e-9)))
    assert el.get_instance_data(g1, g1, wl * (1 + 3e-9)) is not a
    assert el.get_instance_data(g3, g3, wl) is not a
    assert len(el.made) == 3
    # a 32-wavelength cycle stays cached with the default size
    wls = wl * np.linspace(0.95, 1.05, 32)
    for _ in range(2):
        for w in wls:
            el.get_instance_data(g1, g1, w)
    assert len(el.made) == 3 + 32
    nw = _Counting(wavelength_dependent=False)
    b = nw.get_instance_data(g1, g1, 1.0)
    assert nw.get_instance_data(g1, g1, 2.0) is b
    ng = _Counting(grid_dependent=False)
    c = ng.get_instance_data(g1, g1, 1.0)
    assert ng.get_instance_data(g3, g3, 1.0) is c
