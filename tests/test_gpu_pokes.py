"""Poke sets (hcipy_poke_set_*): focal intensities of every single-actuator DM poke for several
wavelengths and pupil sizes in one set of grouped launches, against the oracle's DM poke
(dm_forward with the poked influence function, hcipy/optics/deformable_mirror.py:200-263) followed
by the Fraunhofer transform (hcipy/propagation/fraunhofer.py:56-71), i.e. the reference's
poke-and-propagate loop.  The windowed evaluation (|P(E) + P_w(dE_a)|^2) must equal the full
propagation of the poked field: 1e-12 in FP64, 5e-7 of the base peak in the split precisions."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import workloads
from oracle import hcipy_oracle as o

from helpers import host, og, say

pytestmark = pytest.mark.gpu


def _oracle_poke_intensity(E, pupil, focal, wl, n_act, pitch, h, acts):
    IF = o.gaussian_influence_functions(og(pupil), n_act, pitch, only=acts)
    out = []
    for row in IF:
        Ea = o.dm_forward(E, h * row, wl)
        out.append(np.abs(o.fraunhofer_forward(Ea, og(pupil), og(focal), wl)) ** 2)
    return np.array(out)


def _small_items(cdt, precision):
    """Pupils 128^2 (16 x 16 actuators, 8 px pitch) and 256^2 (16 x 16, 16 px pitch) at three
    wavelengths each; windows narrower than the pupils, clipped at the edges."""
    rng = np.random.default_rng(21)
    lam0 = 500e-9
    items, oracle_in = [], []
    for n, n_act in ((128, 16), (256, 16)):
        pupil = ha.make_pupil_grid(n, 1.0)
        focal = ha.make_focal_grid(4, 8, spatial_resolution=lam0)
        x, y = pupil.x, pupil.y
        E = ((x ** 2 + y ** 2) <= 0.25) * np.exp(0.05j * rng.standard_normal(pupil.size))
        IF = ha.make_gaussian_influence_functions(pupil, n_act, 1.0 / n_act)
        for wl in lam0 * np.array([0.95, 1.0, 1.05]):
            prop = ha.FraunhoferPropagator(pupil, focal, precision=precision)
            items.append((prop, wl, torch.as_tensor(E, device='cuda').to(cdt), IF))
            oracle_in.append((E.astype(np.complex64).astype(np.complex128) if cdt == torch.complex64 else E,
                              pupil, focal, wl, n_act))
    return items, oracle_in, 1e-3 / (2 * 2 * np.pi / lam0)


@pytest.mark.parametrize('precision,cdt,tol', [(None, torch.complex128, 1e-12), ('f16x2', torch.complex64, 5e-7),
                                               ('bf16x3', torch.complex64, 5e-7), ('fp32', torch.complex64, 3e-6)])
def test_poke_set_matches_oracle_pokes(precision, cdt, tol):
    items, oin, h = _small_items(cdt, precision)
    ps = ha.PokeSet(items, h)
    base, pokes = ps.execute()
    torch.cuda.synchronize()
    acts = [0, 7, 15, 120, 136, 255]  # corners and edges (clipped windows), centre
    worst = 0.0
    for k, (E, pupil, focal, wl, n_act) in enumerate(oin):
        b_ref = np.abs(o.fraunhofer_forward(E, og(pupil), og(focal), wl)) ** 2
        peak = b_ref.max()
        eb = np.abs(host(base[k]) - b_ref).max() / peak
        ref = _oracle_poke_intensity(E, pupil, focal, wl, n_act, 1.0 / n_act, h, acts)
        got = host(pokes[k])[acts]
        ep = np.abs(got - ref).max() / peak
        # the poke must change the image (the check is not the base image twice)
        assert np.abs(ref - b_ref).max() > 1e-6 * peak
        worst = max(worst, eb, ep)
        assert eb < tol and ep < tol, (k, eb, ep)
    say('poke set', precision, 'worst peak-relative error %.3e' % worst)


@pytest.mark.parametrize('precision', ['f16x2', 'bf16x3'])
def test_poke_set_every_tile_shape(precision):
    """Every tile shape the poke set's launches can take (each on both stages) gives the oracle's
    pokes; shapes of other precisions are refused."""
    from hcipy_amd import _lib
    items, oin, h = _small_items(torch.complex64, precision)
    items, oin = items[3:4], oin[3:4]  # the 256^2 pupil at the first wavelength
    E, pupil, focal, wl, n_act = oin[0]
    acts = [0, 17, 136, 255]
    b_ref = np.abs(o.fraunhofer_forward(E, og(pupil), og(focal), wl)) ** 2
    ref = _oracle_poke_intensity(E, pupil, focal, wl, n_act, 1.0 / n_act, h, acts)
    for cfg in _lib.POKE_TILE_CFGS[precision]:
        ps = ha.PokeSet(items, h, tiles=(cfg, cfg))
        base, pokes = ps.execute()
        torch.cuda.synchronize()
        peak = b_ref.max()
        eb = np.abs(host(base[0]) - b_ref).max() / peak
        ep = np.abs(host(pokes[0])[acts] - ref).max() / peak
        say('poke set', precision, 'tiles', cfg, 'errors %.3e %.3e' % (eb, ep))
        assert eb < 5e-7 and ep < 5e-7, (cfg, eb, ep)
        ps.close()
    other = [c for c in _lib.POKE_TILE_CFGS['f16x2'] if c not in _lib.POKE_TILE_CFGS[precision]]
    if other:
        with pytest.raises(_lib.HcipyError):
            ha.PokeSet(items, h, tiles=(other[0], other[0]))


def test_poke_set_reexecutes_on_updated_fields_and_replays_in_a_graph():
    items, oin, h = _small_items(torch.complex64, 'f16x2')
    ps = ha.PokeSet(items, h, autotune=True)
    say('autotuned poke-set tile shapes', ps.tiles)
    b1 = [t.clone() for t in ps.execute()[1]]
    # a changed base field changes every result; the same field gives bit-identical results
    saved = [f.clone() for f in ps.fields]
    for f in ps.fields:
        f.mul_(torch.exp(1j * torch.linspace(0, 1, f.numel(), device='cuda')).to(f.dtype))
    b2 = [t.clone() for t in ps.execute()[1]]
    assert all(not torch.equal(a, b) for a, b in zip(b1, b2))
    for f, s in zip(ps.fields, saved):
        f.copy_(s)
    b3 = [t.clone() for t in ps.execute()[1]]
    assert all(torch.equal(a, b) for a, b in zip(b1, b3))
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ps.execute()
    for r in range(3):
        for t in ps.pokes:
            t.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(b1, ps.pokes)), r


def test_poke_set_broadband_jacobian_batch_full_size():
    """32 wavelengths x 1024 pokes (32 x 32 DM) over 256^2 / 512^2 / 1024^2 pupils, complex64,
    split FP16: sampled instances of every pupil size against the oracle, 5e-7 of the base peak."""
    wl_items, h, n_act = workloads.poke_jacobian()
    IFs, props, items = {}, {}, []
    for pupil, focal, wl, E in wl_items:
        n = pupil.shape[0]
        if n not in IFs:
            IFs[n] = ha.make_gaussian_influence_functions(pupil, n_act, 1.0 / n_act)
            props[n] = ha.FraunhoferPropagator(pupil, focal, precision='f16x2')
        items.append((props[n], wl, torch.as_tensor(E.astype(np.complex64), device='cuda'), IFs[n]))
    ps = ha.PokeSet(items, h)
    say('poke set of %d pokes, %.1f GB of device memory' % (sum(p.shape[0] for p in ps.pokes), ps.nbytes / 1e9))
    base, pokes = ps.execute()
    torch.cuda.synchronize()
    for k in (0, 1, 2, 31):
        pupil, focal, wl, E = wl_items[k]
        E64 = E.astype(np.complex64).astype(np.complex128)
        b_ref = np.abs(o.fraunhofer_forward(E64, og(pupil), og(focal), wl)) ** 2
        peak = b_ref.max()
        acts = [0, 527, 1023] if k != 31 else [33, 496]
        ref = _oracle_poke_intensity(E64, pupil, focal, wl, n_act, 1.0 / n_act, h, acts)
        ep = np.abs(host(pokes[k])[acts] - ref).max() / peak
        eb = np.abs(host(base[k]) - b_ref).max() / peak
        say('  item %d (%d^2 pupil, %.1f nm): base %.2e, pokes %.2e' % (k, pupil.shape[0], wl * 1e9, eb, ep))
        assert eb < 5e-7 and ep < 5e-7
