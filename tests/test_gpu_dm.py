"""Deformable mirror (matrix-free Gaussian and Xinetics influence functions) and the
region-of-interest mean against the FP64 oracle (GPU).

Tolerances: surface and its adjoint 1e-12 of the peak (FP64 on both sides, different summation
order); ROI mean 1e-12 relative.
"""
import numpy as np
import pytest
import torch

import os

import hcipy_amd as ha
from oracle import hcipy_oracle as o

from helpers import dev, host, og, say

pytestmark = pytest.mark.gpu

CASES = [
    # pupil px, extent, n_act, tilts (x, y, z), oversampling
    (64, 1.0, 8, (0.0, 0.0, 0.0), 1),
    (96, 1.5, 12, (np.radians(15), np.radians(10), np.radians(20)), 1),
    (80, 1.2, 10, (np.radians(5), np.radians(-8), np.radians(-30)), 3),
    # no z tilt: the separable kernels (per-axis Gaussian factors), with x / y tilts and dithers
    (128, 1.1, 16, (np.radians(12), np.radians(-7), 0.0), 2),
    (192, 1.0, 24, (0.0, 0.0, 0.0), 1),
    # two 256-column blocks and a partial block of 8 rows in the surface kernels
    (268, 1.0, 10, (np.radians(6), 0.0, 0.0), 2),
]


@pytest.mark.parametrize('npix,extent,n_act,tilts,ovs', CASES)
def test_surface_and_adjoint_vs_oracle(npix, extent, n_act, tilts, ovs):
    grid = ha.make_pupil_grid(npix, extent)
    pitch = 1.0 / n_act
    IFo = o.gaussian_influence_functions(og(grid), n_act, pitch, x_tilt=tilts[0], y_tilt=tilts[1], z_tilt=tilts[2],
                                         oversampling=ovs)
    IF = ha.make_gaussian_influence_functions(grid, n_act, pitch, x_tilt=tilts[0], y_tilt=tilts[1], z_tilt=tilts[2],
                                              oversampling=ovs)
    rng = np.random.default_rng(7)
    a = rng.standard_normal((2, n_act * n_act)) * 1e-8
    say('dm', npix, n_act, tilts, ovs)
    s = host(IF.linear_combination(a))
    s_ref = a @ IFo
    assert np.abs(s - s_ref).max() <= 1e-12 * np.abs(s_ref).max()
    g = rng.standard_normal((2, grid.size))
    ga = host(IF.linear_combination_adjoint(dev(g)))
    ga_ref = g @ IFo.T
    assert np.abs(ga - ga_ref).max() <= 1e-12 * np.abs(ga_ref).max()
    # adjoint identity <IF a, g> = <a, IF^T g>
    lhs = np.sum(s * g)
    rhs = np.sum(a * ga)
    assert abs(lhs - rhs) <= 1e-12 * np.sum(np.abs(s) * np.abs(g))


XINETICS_TABLE = np.squeeze(o.read_fits_image(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                          'hcipy_amd', 'data', 'influence_dm5v2.fits')))


@pytest.mark.parametrize('npix,extent,n_act,tilts', [(64, 1.1, 8, (0.0, 0.0, 0.0)),
                                                     (96, 1.5, 12, (np.radians(15), np.radians(10), np.radians(20)))])
def test_xinetics_surface_and_adjoint_vs_oracle(npix, extent, n_act, tilts):
    grid = ha.make_pupil_grid(npix, extent)
    pitch = 1.0 / n_act
    IFo = o.xinetics_influence_functions(og(grid), n_act, pitch, XINETICS_TABLE, *tilts)
    IF = ha.make_xinetics_influence_functions(grid, n_act, pitch, x_tilt=tilts[0], y_tilt=tilts[1], z_tilt=tilts[2])
    rng = np.random.default_rng(8)
    a = rng.standard_normal((2, n_act * n_act)) * 1e-8
    say('xinetics dm', npix, n_act, tilts)
    s = host(IF.linear_combination(a))
    s_ref = a @ IFo
    assert np.abs(s - s_ref).max() <= 1e-12 * np.abs(s_ref).max()
    g = rng.standard_normal((2, grid.size))
    ga = host(IF.linear_combination_adjoint(dev(g)))
    ga_ref = g @ IFo.T
    assert np.abs(ga - ga_ref).max() <= 1e-12 * np.abs(ga_ref).max()
    lhs = np.sum(s * g)
    rhs = np.sum(a * ga)
    assert abs(lhs - rhs) <= 1e-12 * np.sum(np.abs(s) * np.abs(g))


@pytest.mark.parametrize('kind', ['gaussian', 'xinetics'])
@pytest.mark.parametrize('n_act', [12, 16])
def test_reference_mirror_behaviour(n_act, kind):
    """The reference's DM test (tests/test_optics.py:119-193) on the GPU mirror, for both
    influence-function families it covers."""
    x_tilt, y_tilt, z_tilt = np.radians(15), np.radians(10), np.radians(20)
    grid = ha.make_pupil_grid(256, 1.5)
    pitch = 1.0 / n_act
    pos = ha.make_actuator_positions(n_act, pitch)
    assert np.allclose(pos.delta, pitch)
    pos_t = ha.make_actuator_positions(n_act, pitch, x_tilt=x_tilt, y_tilt=y_tilt, z_tilt=z_tilt)
    make = ha.make_gaussian_influence_functions if kind == 'gaussian' else ha.make_xinetics_influence_functions
    IF = make(grid, n_act, pitch, x_tilt=x_tilt, y_tilt=y_tilt, z_tilt=z_tilt)
    dm = ha.DeformableMirror(IF)
    assert len(IF) == n_act ** 2 and dm.num_actuators == n_act ** 2
    assert IF.is_sparse
    x = torch.as_tensor(grid.x, device='cuda')
    y = torch.as_tensor(grid.y, device='cuda')
    for act, p in zip(IF, pos_t.points):
        t = act.as_subclass(torch.Tensor)
        c = [((t * x).sum() / t.sum()).item(), ((t * y).sum() / t.sum()).item()]
        assert np.allclose(c, p, atol=0.02)
    assert host(dm.surface).std() < 1e-12
    for i in np.random.default_rng(1).integers(0, n_act ** 2, size=5):
        dm.actuators[i] = np.random.default_rng(int(i)).standard_normal()
        assert host(dm.surface).std() > 1e-12
        dm.actuators[i] = 0
    assert host(dm.surface).std() < 1e-12
    dm.random(0.001)
    assert host(dm.surface).std() > 1e-12
    wf = ha.Wavefront(ha.Field(torch.ones(grid.size, dtype=torch.complex128, device='cuda'), grid))
    wf_out = dm.forward(wf)
    assert np.allclose(np.angle(host(wf_out.electric_field)), np.angle(np.exp(1j * host(dm.phase_for(1)))))
    wf_in = dm.backward(wf_out)
    assert np.allclose(host(wf_in.electric_field), host(wf.electric_field))
    assert np.allclose(host(dm.opd), 2 * host(dm.surface))
    dm.flatten()
    assert host(dm.surface).std() < 1e-12


def test_mirror_forward_vs_oracle():
    grid = ha.make_pupil_grid(64)
    n_act = 8
    IFo = o.gaussian_influence_functions(og(grid), n_act, 1.0 / n_act)
    dm = ha.DeformableMirror(ha.make_gaussian_influence_functions(grid, n_act, 1.0 / n_act))
    dm.actuators = np.random.default_rng(3).standard_normal(n_act ** 2) * 5e-8
    E = np.exp(1j * np.linspace(0, 1, grid.size))
    wl = 6e-7
    out = host(dm.forward(ha.Wavefront(ha.Field(dev(E), grid), wl)).electric_field)
    ref = o.dm_forward(E, o.dm_surface(IFo, dm.actuators), wl)
    assert np.abs(out - ref).max() < 1e-11


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_mean_intensity_in_roi(dtype):
    rng = np.random.default_rng(5)
    img = rng.random((3, 1000)).astype(dtype)
    mask = rng.random(1000) > 0.6
    got = host(ha.get_mean_intensity_in_roi(dev(img), dev(mask)))
    ref = np.array([o.mean_intensity_in_roi(img[b].astype(np.float64), mask) for b in range(3)])
    np.testing.assert_allclose(got, ref, rtol=1e-12)
    single = ha.get_mean_intensity_in_roi(dev(img[1]), dev(mask.astype(np.float64)))
    assert abs(single - ref[1]) <= 1e-12 * abs(ref[1])
    empty = ha.get_mean_intensity_in_roi(dev(img[0]), dev(np.zeros(1000, dtype=bool)))
    assert np.isnan(empty)
