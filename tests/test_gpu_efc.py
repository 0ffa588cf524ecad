"""Electric-field-conjugation pieces on the device: the Tikhonov-regularised inverse
(hcipy/util/matrix_inversion.py:71-100) against the oracle's SVD restatement, and one EFC step
built from a poke-set Jacobian (the DM-poke loop of the reference's EFC tutorial,
doc/tutorial_notebooks/ElectricFieldConjugation) that must match the oracle's poke differences
and deepen the dark hole."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from oracle import hcipy_oracle as o

from helpers import host, og, say

pytestmark = pytest.mark.gpu


def _matrix(m, n, decay, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (U * np.logspace(0, -decay, n)) @ V.T * 3.7


@pytest.mark.parametrize('m,n,rcond,tol', [(601, 200, 1e-2, 1e-10), (601, 200, 1e-3, 1e-8), (130, 64, 0.3, 1e-12),
                                           (4000, 1024, 1e-2, 1e-10)])
def test_inverse_tikhonov_matches_oracle(m, n, rcond, tol):
    M = _matrix(m, n, 4, n + m)
    ref = o.inverse_tikhonov(M, rcond)
    got = ha.inverse_tikhonov(torch.as_tensor(M, device='cuda'), rcond)
    err = np.abs(host(got) - ref).max() / np.abs(ref).max()
    s_max = np.linalg.svd(M, compute_uv=False).max()
    say('inverse_tikhonov %d x %d rcond %g: %.3e (s_max %.3e)' % (m, n, rcond, err, abs(got.s_max / s_max - 1)))
    assert got.shape == (n, m)
    assert abs(got.s_max / s_max - 1) < 1e-13
    assert err < tol


def test_inverse_tikhonov_rejects_bad_arguments():
    with pytest.raises(ValueError):
        ha.inverse_tikhonov(torch.zeros(3, 4, 5, device='cuda'))
    with pytest.raises(ValueError):
        ha.inverse_tikhonov(torch.zeros(3, 4, dtype=torch.complex128, device='cuda'))
    with pytest.raises(ha._lib.HcipyError):
        ha.inverse_tikhonov(torch.zeros(6, 4, device='cuda'), 1e-2)


def test_efc_step_from_poke_set_jacobian():
    """FP64: 128^2 pupil with a 0.1 rad phase aberration, 16 x 16 Gaussian DM, focal grid q = 4,
    8 lambda/D, D-shaped dark zone 2 < r < 6 lambda/D, x > 1.  Jacobian = field change of every
    poke / poke height (poke set), control = -inverse_tikhonov([Re J; Im J], 1e-2) [Re dE; Im dE]
    with dE the aberration's field in the zone (the aberrated minus the unaberrated focal field:
    without a coronagraph the Airy rings are not the mirror's to remove).  One step must remove
    more than 98 % of the aberration's light there (the oracle's own step removes 99.8 %)."""
    lam = 500e-9
    pupil = ha.make_pupil_grid(128, 1.0)
    focal = ha.make_focal_grid(4, 8, spatial_resolution=lam)
    ap = ((pupil.x ** 2 + pupil.y ** 2) <= 0.25).astype(float)
    from hcipy_amd import workloads
    E = ap * np.exp(1j * workloads.smoothed_phase(pupil.shape, 3.0, 0.1, ap, 9))
    n_act = 16
    IF = ha.make_gaussian_influence_functions(pupil, n_act, 1.0 / n_act)
    prop = ha.FraunhoferPropagator(pupil, focal)
    h = 1e-10
    ps = ha.PokeSet([(prop, lam, torch.as_tensor(E, device='cuda'), IF)], h, output='field_change')
    base, dfield = ps.execute()
    xf, yf = focal.x / lam, focal.y / lam
    r = np.hypot(xf, yf)
    dz = (r > 2) & (r < 6) & (xf > 1)
    J = host(dfield[0])[:, dz].T / h  # (N_dz, n_act) complex
    # the Jacobian against the oracle's poke-and-propagate differences on sampled actuators
    E0 = o.fraunhofer_forward(E, og(pupil), og(focal), lam)
    acts = [40, 119, 136, 200]
    IFo = o.gaussian_influence_functions(og(pupil), n_act, 1.0 / n_act, only=acts)
    Jo = np.array([(o.fraunhofer_forward(o.dm_forward(E, h * row, lam), og(pupil), og(focal), lam) - E0)[dz] / h
                   for row in IFo]).T
    ej = np.abs(J[:, acts] - Jo).max() / np.abs(Jo).max()
    say('Jacobian vs oracle poke differences: %.3e' % ej)
    assert ej < 1e-7  # the oracle's differences cancel 7 digits of a 1e-3 rad poke
    # control matrix and one EFC step
    G = np.concatenate([J.real, J.imag])
    Gi = ha.inverse_tikhonov(torch.as_tensor(G, device='cuda'), 1e-2)
    ref = o.inverse_tikhonov(G, 1e-2)
    eg = np.abs(host(Gi) - ref).max() / np.abs(ref).max()
    say('EFC control matrix vs oracle SVD inverse: %.3e' % eg)
    assert eg < 1e-10
    Ep = o.fraunhofer_forward(ap.astype(complex), og(pupil), og(focal), lam)
    dE = (E0 - Ep)[dz]
    da = -host(Gi) @ np.concatenate([dE.real, dE.imag])
    dm = ha.DeformableMirror(IF)
    dm.actuators = da
    wf = dm.forward(ha.Wavefront(ha.Field(torch.as_tensor(E, device='cuda'), pupil), lam))
    E1 = host(prop.forward(wf).electric_field)
    before = (np.abs(dE) ** 2).mean()
    after = (np.abs((E1 - Ep)[dz]) ** 2).mean()
    say('aberration light in the dark zone: %.3e -> %.3e' % (before, after))
    assert after < before / 50
