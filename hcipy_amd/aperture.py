"""Aperture field generators (hcipy/aperture/generic.py:11-46, 301-348, 394-429, 870-895).

Generators are evaluated on the host in FP64 (set-up work) and return Fields on the default
device.
"""
import numpy as np

from .field import Field

__all__ = ['make_circular_aperture', 'make_rectangular_aperture', 'make_spider', 'make_obstructed_circular_aperture',
           'make_super_gaussian_aperture']


def _xy(grid):
    if grid.is_separated:
        x, y = grid.separated_coords
        return x[np.newaxis, :], y[:, np.newaxis]
    return grid.coords


def make_circular_aperture(diameter, center=None):
    shift = np.zeros(2) if center is None else center * np.ones(2)

    def func(grid):
        x, y = _xy(grid)
        f = (((x - shift[0]) ** 2 + (y - shift[1]) ** 2) <= (diameter / 2) ** 2).ravel()
        return Field(f.astype('float'), grid)
    return func


def make_rectangular_aperture(size=1, center=None):
    size = size * np.ones(2)
    shift = np.zeros(2) if center is None else center * np.ones(2)

    def func(grid):
        x, y = _xy(grid)
        f = ((np.abs(x - shift[0]) <= size[0] / 2) * (np.abs(y - shift[1]) <= size[1] / 2)).ravel()
        return Field(f.astype('float'), grid)
    return func


def _spider_mask(grid, p1, p2, spider_width):
    delta = np.array(p2) - np.array(p1)
    shift = delta / 2 + np.array(p1)
    spider_angle = np.arctan2(delta[1], delta[0])
    spider_length = np.linalg.norm(delta)
    x, y = _xy(grid)
    x = x - shift[0]
    y = y - shift[1]
    x_new = x * np.cos(spider_angle) + y * np.sin(spider_angle)
    y_new = y * np.cos(spider_angle) - x * np.sin(spider_angle)
    spider = x_new <= (spider_length / 2)
    spider = spider * (x_new >= (-spider_length / 2))
    spider = spider * (y_new <= (spider_width / 2))
    spider = spider * (y_new >= (-spider_width / 2))
    return (1 - spider.ravel()).astype(float)


def make_spider(p1, p2, spider_width):
    def func(grid):
        return Field(_spider_mask(grid, p1, p2, spider_width), grid)
    return func


def make_obstructed_circular_aperture(pupil_diameter, central_obscuration_ratio, num_spiders=0, spider_width=0.01):
    def func(grid):
        x, y = _xy(grid)
        outer = ((x ** 2 + y ** 2) <= (pupil_diameter / 2) ** 2).ravel().astype(float)
        d_in = pupil_diameter * central_obscuration_ratio
        inner = ((x ** 2 + y ** 2) <= (d_in / 2) ** 2).ravel().astype(float)
        spiders = 1
        for angle in np.linspace(0, 2 * np.pi, num_spiders, endpoint=False):
            spiders = spiders * _spider_mask(grid, (0, 0), (pupil_diameter * np.cos(angle), pupil_diameter * np.sin(angle)), spider_width)
        return Field((outer - inner) * spiders, grid)
    return func


def make_super_gaussian_aperture(full_width_half_maximum, shape_parameter=5, center=None):
    """exp(-(r / sigma)^(2 p) / 2) with sigma = fwhm / 2 / (2 ln 2)^(1 / (2 p)), so that the value
    is 1/2 at r = fwhm / 2; p = 1 is a Gaussian (generic.py:870-895)."""
    shift = np.zeros(2) if center is None else center * np.ones(2)

    def func(grid):
        x, y = _xy(grid)
        sigma = 0.5 * full_width_half_maximum / (2 * np.log(2)) ** (1 / (2 * shape_parameter))
        r = np.hypot(x - shift[0], y - shift[1]) / sigma
        return Field(np.exp(-0.5 * np.abs(r) ** (2 * shape_parameter)).ravel(), grid)
    return func
