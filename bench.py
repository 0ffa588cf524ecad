#!/usr/bin/env python3
"""Throughput of the MFMA matrix Fourier transform: FraunhoferPropagator forward to the focal
intensity, 1024^2 pupil -> 512^2 focal plane (q = 4, 64 lambda/D), complex64 input, batch 1.

One step = one replay of a HIP graph holding `--chunk` propagations (each: data split + two
MFMA GEMM stages, |E|^2 formed in the epilogue) of a field already resident in HBM, so that the
driver's short runs still time tens of milliseconds of GPU work; `value` is propagations/s.
Also reported: the same propagation with complex64 / complex128 field output and the plain-FP32
MFMA path, each with its peak-relative intensity error against this library's FP64 MFMA path
and, in the CPU leg, against the FP64 oracle; the dominant kernel's roofline from HIP events;
the other paths of the hot path (angular spectrum, zoom FFT, Lyot chain with its error, vortex,
Jones filter, FFT, pointwise and deformable-mirror kernels in GB/s); and the CPU baseline
(numpy/scipy restatement of the reference's two cgemm calls, timed on the host cores).
"""
import argparse
import math
import ctypes
import json
import os
import sys
import time

METRIC = 'MFT propagations/s 1024²→512² (split-prec MFMA vs FP32) at ≤1e-7 peak-rel err'
PEAK_TFLOPS = {'f16x2': 2500.0, 'bf16x3': 2500.0, 'fp32': 157.3, 'fp64': 78.6}  # MI355X dense MFMA peaks


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200, help='graph replays timed (each --chunk propagations)')
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--precision', default='f16x2', choices=['f16x2', 'bf16x3', 'fp32'])
    ap.add_argument('--chunk', type=int, default=64, help='propagations per step (one captured HIP graph)')
    ap.add_argument('--no-graph', action='store_true')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-autotune', action='store_true', help='keep the default tile shapes')
    ap.add_argument('--no-variants', action='store_true', help='time only the main path (for counter runs)')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run a plain 256 MiB device-to-device copy 3 times (scale of the HBM byte counters)')
    ap.add_argument('--cpu-seconds', type=float, default=2.0, help='t_max of the CPU timing loop')
    ap.add_argument('--paths-only', action='store_true',
                    help='only time the angular-spectrum, Lyot-chain, vortex and mixed-size grouped paths')
    ap.add_argument('--path', default=None, choices=['poke_jacobian', 'mixed_grouped', 'broadband_grouped', 'fft_2048',
                                                     'darkhole_gradient'],
                    help='only time this one workload (for profiler runs)')
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU: the MFT path has no multi-GPU mode in this build '
              '(--gpus %d requested).' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_read': src.numel() * 4, 'copy_bytes_written': dst.numel() * 4, 'copies': 3}))
        return

    import hcipy_amd as ha
    from hcipy_amd import _lib, workloads
    pupil, focal, E = workloads.plain_mft_1024()
    E64 = E.astype(np.complex64)
    x = torch.as_tensor(E64, device=dev)
    wavelength = 1.0

    tiles = {}
    forms = {}
    lyot_outs = {}

    def setup(precision, parity=True):
        prop = ha.FraunhoferPropagator(pupil, focal, precision=precision)
        d = prop.instance(wavelength)
        mft = d.fourier_transform
        alpha = complex(pupil.weights) * d.norm_factor
        cdt = torch.complex64 if precision != 'fp64' else torch.complex128
        plan = mft.plan_for(cdt, dev)
        if not parity and plan.parity_available:
            plan.parity = False  # the complex form, for comparison
        forms[(precision, parity)] = 'parity' if plan.parity_available and plan.parity else 'complex'
        if not args.no_autotune:
            tiles[(precision, forms[(precision, parity)])] = mft.autotune(ha.Field(x, pupil), repeats=10)['forward']
        ws = mft._workspace_for(plan, 1, dev)
        return prop, plan, ws, alpha

    def time_path(precision, steps, warmup, output, parity=True):
        """Seconds per propagation over `steps` steps of `chunk` propagations each (one graph
        replay per step), after `warmup` untimed steps; HIP-event durations of the three steps
        (split, stage 1, stage 2 with the parity recombination)."""
        prop, plan, ws, alpha = setup(precision, parity)
        odt = {_lib.OUTPUT_FIELD: torch.complex64, _lib.OUTPUT_INTENSITY: torch.float32,
               _lib.OUTPUT_FIELD_C128: torch.complex128}[output]
        out = torch.empty(focal.size, dtype=odt, device=dev)
        chunk = max(1, args.chunk)

        def prop_once():
            plan.execute(_lib.FORWARD, x, out, 1, alpha, ws, output)

        prop_once()
        torch.cuda.synchronize()
        if args.no_graph:
            def step():
                for _ in range(chunk):
                    prop_once()
        else:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                prop_once()
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                for _ in range(chunk):
                    prop_once()
            step = graph.replay
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # per-kernel durations from HIP events on the launching stream: each step inside whole
        # propagations (events between the launches of one propagation, 200 propagations), and
        # each kernel alone over 200 back-to-back launches
        kms = (ctypes.c_float * 3)()
        st = _lib.load().hcipy_mft_profile_pipeline(plan.handle, _lib.FORWARD, _lib.dptr(x), _lib.dptr(out), 1,
                                                    float(alpha.real), float(alpha.imag), output, _lib.dptr(ws),
                                                    ws.numel(), 200, kms, _lib.stream_ptr(dev))
        _lib.check(st, 'hcipy_mft_profile_pipeline')
        kiso = (ctypes.c_float * 3)()
        st = _lib.load().hcipy_mft_profile(plan.handle, _lib.FORWARD, _lib.dptr(x), _lib.dptr(out), 1,
                                           float(alpha.real), float(alpha.imag), output, _lib.dptr(ws), ws.numel(),
                                           200, kiso, _lib.stream_ptr(dev))
        _lib.check(st, 'hcipy_mft_profile')
        time_path.isolated[(precision, output, parity)] = [kiso[0], kiso[1], kiso[2]]
        return (t1 - t0) / (steps * chunk), [kms[0], kms[1], kms[2]], out.clone()

    # Reference intensity: this library's FP64 MFMA path on the same complex64-rounded input.
    prop64 = ha.FraunhoferPropagator(pupil, focal)
    wf64 = ha.Wavefront(ha.Field(torch.as_tensor(E64.astype(np.complex128), device=dev), pupil), wavelength)
    I_ref = prop64.forward_intensity(wf64).as_subclass(torch.Tensor)
    peak = I_ref.max().item()

    def err(o):
        I = o.to(torch.float64) if not o.is_complex() else o.to(torch.complex128).abs() ** 2
        return (I - I_ref).abs().max().item() / peak

    def time_broadband(n_wl=32, batch=32, reps=20):
        """Broadband / poke-batch workload: 32 wavelengths over a 10 % band (per-wavelength
        Fourier matrices, focal grid fixed in lambda0/D), 256^2 pupil -> 128^2 focal plane,
        a batch of 32 fields per wavelength, f16x2, focal intensity.  Grouped launch vs one
        call per wavelength."""
        pg = ha.make_pupil_grid(256)
        fg = ha.make_focal_grid(4, 16)
        prop = ha.FraunhoferPropagator(pg, fg, precision='f16x2')
        rng = np.random.default_rng(4)
        Eb = torch.as_tensor((rng.standard_normal((batch, pg.size)) + 1j * rng.standard_normal((batch, pg.size)))
                             .astype(np.complex64), device=dev)
        wls = np.linspace(0.95, 1.05, n_wl)
        wfs = [ha.Wavefront(ha.Field(Eb, pg), wl) for wl in wls]
        grp = ha.MftGroup(n_wl)
        prop.forward_multiple(wfs, output='intensity', group=grp)
        for wf in wfs:
            prop.forward_intensity(wf)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            prop.forward_multiple(wfs, output='intensity', group=grp)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            for wf in wfs:
                prop.forward_intensity(wf)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        n = n_wl * batch * reps
        return {'workload': '32 wavelengths x 32 fields, 256^2 -> 128^2, f16x2 intensity',
                'grouped_propagations_per_s': round(n / (t1 - t0), 1),
                'per_wavelength_calls_propagations_per_s': round(n / (t2 - t1), 1)}

    def time_graph(fn, n_chunks=10, chunk=10):
        """Seconds per call of ``fn`` replayed from a captured graph of ``chunk`` calls (inputs
        resident in HBM), and the outputs of the last captured call."""
        fn()
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(chunk):
                outs = fn()
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n_chunks):
            g.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / (n_chunks * chunk), outs

    def time_fft_2048():
        """FastFourierTransform of a 2048^2 field (q = 1) on the direct power-of-two FFT."""
        out = {}
        rng = np.random.default_rng(8)
        # A row pass (field in, intermediate out) and a column pass (intermediate in, field out);
        # the intermediate is FP64 complex except for precision 'fp32'.
        fgrid = ha.make_pupil_grid(2048, 1.0)
        fft = ha.FastFourierTransform(fgrid)
        fft32 = ha.FastFourierTransform(fgrid, precision='fp32')
        fx = torch.as_tensor(rng.standard_normal(fgrid.size) + 1j * rng.standard_normal(fgrid.size), device=dev)
        fref = ha.MatrixFourierTransform(fgrid, fft.output_grid).forward(ha.Field(fx, fgrid)).as_subclass(torch.Tensor)
        for key, t, xin, ib in (('fft_2048_c64', fft, fx.to(torch.complex64), 16),
                                ('fft_2048_c64_fp32', fft32, fx.to(torch.complex64), 8),
                                ('fft_2048_c128', fft, fx, 16)):
            fin = ha.Field(xin, fgrid)
            sec, o = time_graph(lambda: t.forward(fin))
            o = o.as_subclass(torch.Tensor).to(torch.complex128)
            nbytes = 2 * xin.element_size() * 2048 ** 2 + 2 * ib * 2048 ** 2
            out[key] = {'value': round(1.0 / sec, 1), 'unit': 'transforms/s', 'ms': round(sec * 1e3, 4),
                        'method': t.method, 'hbm_gbs_pass_bytes': round(nbytes / sec / 1e9, 1),
                        'peak_rel_field_error_vs_fp64_mft': (o - fref).abs().max().item() / fref.abs().max().item()}
        return out

    def time_other_paths():
        """The other workloads of the hot path, each timed from captured graphs."""
        out = {}
        # Angular-spectrum propagation of the 1024^2 field: lambda = 500 nm, D = 1 m, z = 1 km,
        # 2x zero padding (2048^2 FFT grid).  Error: complex64 field against the complex128 path.
        asp_grid = ha.make_pupil_grid(1024, 1.0)
        asp = ha.AngularSpectrumPropagator(asp_grid, 1000.0)
        e64 = ha.Field(x, asp_grid)
        e128 = ha.Field(x.to(torch.complex128), asp_grid)
        sec, o = time_graph(lambda: asp.forward(ha.Wavefront(e64, 500e-9)).electric_field)
        ref = asp.forward(ha.Wavefront(e128, 500e-9)).electric_field.as_subclass(torch.Tensor)
        o = o.as_subclass(torch.Tensor).to(torch.complex128)
        NX = 2048
        compulsory = (8 + 8 + 32) * 2 ** 20           # field in + out (c64), transfer function (c64 2048^2)
        passes = (8 + 16 + 16 + 32 + 16 + 16 + 8) * 2 ** 20   # rows, columns (+ transfer function), rows
        out['angular_spectrum_1024_q2_c64'] = {
            'value': round(1.0 / sec, 1), 'unit': 'propagations/s', 'ms': round(sec * 1e3, 4),
            'peak_rel_field_error_vs_c128': (o - ref).abs().max().item() / ref.abs().max().item(),
            'hbm_gbs_compulsory': round(compulsory / sec / 1e9, 1), 'hbm_gbs_pass_bytes': round(passes / sec / 1e9, 1),
            'peak_gbs': 8000.0, 'internal_grid': [NX, NX],
            'gflop': round(2 * 5 * NX * NX * math.log2(NX * NX) / 1e9, 3)}
        # Angular-spectrum propagation of a 4096^2 field (8192^2 FFT grid: 8192-point row and column
        # passes), complex64, same optics; error against the complex128 path.
        big_grid = ha.make_pupil_grid(4096, 1.0)
        big = ha.AngularSpectrumPropagator(big_grid, 1000.0)
        bx = torch.as_tensor(np.random.default_rng(6).standard_normal(big_grid.size), device=dev)
        b64 = ha.Field(bx.to(torch.complex64), big_grid)
        sec, o = time_graph(lambda: big.forward(ha.Wavefront(b64, 500e-9)).electric_field, n_chunks=4, chunk=5)
        bref = big.forward(ha.Wavefront(ha.Field(bx.to(torch.complex128), big_grid), 500e-9)).electric_field
        bref = bref.as_subclass(torch.Tensor)
        o = o.as_subclass(torch.Tensor).to(torch.complex128)
        out['angular_spectrum_4096_q2_c64'] = {
            'value': round(1.0 / sec, 1), 'unit': 'propagations/s', 'ms': round(sec * 1e3, 4),
            'peak_rel_field_error_vs_c128': (o - bref).abs().max().item() / bref.abs().max().item(),
            'hbm_gbs_pass_bytes': round(16 * passes / sec / 1e9, 1), 'internal_grid': [8192, 8192]}
        del bref, o, b64, bx
        # The same 1024^2 -> 512^2 Fraunhofer transform as a zoom FFT (Bluestein chirp-z per axis):
        # complex64 fields with FP64 chirps, line FFTs and intermediate, and complex128 fields.
        d64 = prop64.instance(wavelength)
        for key, prec, xin in (('zoom_fft_1024_to_512_c64_fp64', 'fp64', x), ('zoom_fft_1024_to_512_c64_fp32', 'fp32', x),
                               ('zoom_fft_1024_to_512_c128', None, x.to(torch.complex128))):
            zf = ha.ZoomFastFourierTransform(pupil, d64.uv_grid, precision=prec)
            zin = ha.Field(xin, pupil)
            sec, o = time_graph(lambda: zf.forward(zin))
            o = o.as_subclass(torch.Tensor).to(torch.complex128) * complex(d64.norm_factor)
            out[key] = {'value': round(1.0 / sec, 1), 'unit': 'transforms/s', 'ms': round(sec * 1e3, 4),
                        'peak_rel_intensity_error': err(o)}
        # Lyot coronagraph chain: 1024^2 pupil -> 256^2 focal-plane mask -> Lyot plane (stop 0.95 D)
        # -> 512^2 focal intensity, split-FP16, plain FP32 and FP64 (complex128); the error of each
        # against the oracle's FP64 chain relative to the unocculted peak is added in the CPU leg.
        lp, lfg, lfocal, lE, lfpm, lstop, _ = workloads.lyot_1024()
        for key, prec, cdt in (('lyot_chain_1024_f16x2', 'f16x2', torch.complex64),
                               ('lyot_chain_1024_bf16x3', 'bf16x3', torch.complex64),
                               ('lyot_chain_1024_fp32', 'fp32', torch.complex64),
                               ('lyot_chain_1024_fp64', None, torch.complex128)):
            coro = ha.LyotCoronagraph(lp, ha.Field(torch.as_tensor(lfpm, device=dev), lfg),
                                      ha.Field(torch.as_tensor(lstop, device=dev), lp), precision=prec)
            lprop = ha.FraunhoferPropagator(lp, lfocal, precision=prec)
            lx = ha.Field(torch.as_tensor(lE.astype(np.complex64), device=dev).to(cdt), lp)
            sec, lo = time_graph(lambda: lprop.forward_intensity(coro.forward(ha.Wavefront(lx, 1.0))),
                                 n_chunks=5 if prec is None else 10)
            lyot_outs[key] = lo.as_subclass(torch.Tensor).clone()
            out[key] = {'value': round(1.0 / sec, 1), 'unit': 'chains/s', 'ms': round(sec * 1e3, 4),
                        'gflop_per_chain': 11.81}
            del coro, lprop
        # HBM-bound kernels of the chain on the 1024^2 complex64 field (apodizer / mask product,
        # phase screen, intensity) and the 48 x 48 deformable-mirror surface: GB/s of the bytes
        # each kernel must read and write.
        amp = torch.as_tensor(np.random.default_rng(8).random(asp_grid.size), device=dev)
        phi = torch.as_tensor(np.random.default_rng(9).standard_normal(asp_grid.size), device=dev)
        xo = torch.empty_like(x)
        n = x.numel()
        kernels = {
            'sub_mul_1024_c64_f64mask': (lambda: ha.pointwise.sub_mul(x, None, amp, out=xo), n * (8 + 8 + 8)),
            'phase_1024_c64_f64phase': (lambda: ha.pointwise.multiply_phase(x, phi, 0.7), n * (8 + 8 + 8)),
            'intensity_1024_c64': (lambda: ha.pointwise.intensity(x), n * (8 + 4)),
        }
        IFb = ha.make_gaussian_influence_functions(asp_grid, 48, 1.0 / 48)
        dmb = ha.DeformableMirror(IFb)
        acts = torch.as_tensor(np.random.default_rng(10).standard_normal(48 * 48) * 1e-9, device=dev)
        kernels['dm_surface_1024_48x48_f64'] = (lambda: IFb.linear_combination(acts), n * 8 + 48 * 48 * 8)
        gsurf = torch.as_tensor(np.random.default_rng(11).standard_normal(n), device=dev)
        kernels['dm_adjoint_1024_48x48_f64'] = (lambda: IFb.linear_combination_adjoint(gsurf), n * 8 + 48 * 48 * 8)
        for key, (fn, nbytes) in kernels.items():
            sec, _ = time_graph(fn, n_chunks=10, chunk=20)
            out[key] = {'us': round(sec * 1e6, 2), 'gbs': round(nbytes / sec / 1e9, 1), 'bytes': nbytes,
                        'frac_of_8tbs': round(nbytes / sec / 8e12, 3)}
        del dmb
        # Charge-2 vortex coronagraph on the 1024^2 pupil: FFT filter (2048^2) + 5 levels of MFTs
        # 1024^2 <-> 128^2 as two grouped launches, f16x2.
        vc = ha.VortexCoronagraph(asp_grid, 2, precision='f16x2')
        sec, _ = time_graph(lambda: vc.forward(ha.Wavefront(e64, 1.0)).electric_field)
        out['vortex_1024_charge2_f16x2'] = {'value': round(1.0 / sec, 1), 'unit': 'forward/s', 'ms': round(sec * 1e3, 4),
                                            'levels': len(vc.focal_grids)}
        # Charge-2 vector vortex on the same pupil: 2 x 2 Jones-matrix field (an unpolarized input).
        vvc = ha.VectorVortexCoronagraph(2, precision='f16x2')
        sec, _ = time_graph(lambda: vvc.forward(ha.Wavefront(e64, 1.0)).electric_field)
        out['vector_vortex_1024_charge2_f16x2'] = {'value': round(1.0 / sec, 1), 'unit': 'forward/s',
                                                   'ms': round(sec * 1e3, 4)}
        # 2 x 2 Jones-matrix transfer function (random unit-modulus elements) on a Jones-vector field
        # of the 1024^2 pupil, q = 2: one forward FFT per input component, one inverse per output.
        rng = np.random.default_rng(5)
        jint = ha.make_fft_grid(asp_grid, 2, 1)
        jtf = torch.as_tensor(np.exp(2j * np.pi * rng.random((2, 2, jint.size))), device=dev)
        jf = ha.FourierFilter(asp_grid, jtf, 2)
        jx128 = ha.Field(torch.stack([x, x.flip(0)]).to(torch.complex128), asp_grid)
        jref = jf.forward(jx128).as_subclass(torch.Tensor)
        for key, xin in (('jones_filter_1024_q2_c64', jx128.as_subclass(torch.Tensor).to(torch.complex64)),
                         ('jones_filter_1024_q2_c128', jx128.as_subclass(torch.Tensor))):
            jin = ha.Field(xin, asp_grid)
            sec, o = time_graph(lambda: jf.forward(jin))
            o = o.as_subclass(torch.Tensor).to(torch.complex128)
            fb = 8 if xin.dtype == torch.complex64 else 16
            # rows: 2 fields in, 2 intermediates out; columns: 2 in, 4 transfer functions, 2 out;
            # rows: 2 in, 2 fields out
            nbytes = fb * (2 * 1024 ** 2 + 2 * 2048 * 1024 + 2 * 2048 * 1024 + 4 * 2048 ** 2 + 2 * 2048 * 1024
                           + 2 * 2048 * 1024 + 2 * 1024 ** 2)
            out[key] = {'value': round(1.0 / sec, 1), 'unit': 'filters/s', 'ms': round(sec * 1e3, 4),
                        'hbm_gbs_pass_bytes': round(nbytes / sec / 1e9, 1),
                        'peak_rel_field_error_vs_c128': (o - jref).abs().max().item() / jref.abs().max().item()}
        out.update(time_fft_2048())
        return out

    def time_mixed_grouped(n_wl=8, batch=4, reps=10):
        """Mixed pupil sizes (256^2, 512^2, 1024^2 -> 128^2, 256^2, 512^2 focal planes, q = 4), each at
        n_wl wavelengths with their own matrices and a batch of fields (pokes), f16x2 intensity:
        every transform as one grouped launch per stage, against one call per transform."""
        rng = np.random.default_rng(6)
        calls, per = [], []
        for npix in (256, 512, 1024):
            pg = ha.make_pupil_grid(npix)
            fg = ha.make_focal_grid(4, npix // 16)
            E = torch.as_tensor((rng.standard_normal((batch, pg.size)) + 1j * rng.standard_normal((batch, pg.size)))
                                .astype(np.complex64), device=dev)
            prop = ha.FraunhoferPropagator(pg, fg, precision='f16x2')
            for wl in np.linspace(0.95, 1.05, n_wl):
                wf = ha.Wavefront(ha.Field(E, pg), wl)
                per.append((prop, wf))
                d = prop.instance(wl)
                calls.append((d.fourier_transform, _lib.FORWARD, E, complex(pg.weights) * d.norm_factor,
                              _lib.OUTPUT_INTENSITY))
        grp = ha.MftGroup(len(calls))
        gflop = sum(8.0 * b * (ny * nx * nu + nv * ny * nu) for b, ny, nx, nv, nu in
                    [(batch, n, n, n // 2, n // 2) for n in (256, 512, 1024) for _ in range(n_wl)]) / 1e9

        def grouped():
            return grp.execute(calls)

        def one_by_one():
            return [prop.forward_intensity(wf) for prop, wf in per]

        sec_g, _ = time_graph(grouped, n_chunks=reps, chunk=1)
        sec_p, _ = time_graph(one_by_one, n_chunks=reps, chunk=1)
        return {'workload': '%d transforms: pupils 256^2/512^2/1024^2 x %d wavelengths x %d fields, f16x2 intensity'
                            % (len(calls), n_wl, batch),
                'grouped_ms': round(sec_g * 1e3, 3), 'one_call_per_transform_ms': round(sec_p * 1e3, 3),
                'grouped_tflops': round(gflop / sec_g / 1e3, 1)}

    poke_sample = {}

    def time_poke_jacobian(reps=5):
        """Broadband DM-poke batch (workloads.poke_jacobian): 32 wavelengths x 1024 single-actuator
        pokes of a 32 x 32 Gaussian DM over 256^2 / 512^2 / 1024^2 pupils -> 128^2 / 256^2 / 512^2
        focal intensities, complex64, split FP16: one poke-set execution (base transforms + windowed
        poke transforms as grouped launches) replayed from a HIP graph."""
        wl_items, h, n_act = workloads.poke_jacobian()
        IFs, props, items = {}, {}, []
        for pg, fg, wl, E in wl_items:
            n = pg.shape[0]
            if n not in IFs:
                IFs[n] = ha.make_gaussian_influence_functions(pg, n_act, 1.0 / n_act)
                props[n] = ha.FraunhoferPropagator(pg, fg, precision='f16x2')
            items.append((props[n], wl, torch.as_tensor(E.astype(np.complex64), device=dev), IFs[n]))
        ps = ha.PokeSet(items, h, autotune=True)
        sec, _ = time_graph(ps.execute, n_chunks=reps, chunk=1)
        k = 2  # a 1024^2 item: one poke kept for the oracle check of the CPU leg
        poke_sample.update(item=wl_items[k], act=527, h=h, n_act=n_act, base=ps.base[k].cpu().numpy(),
                           poke=ps.pokes[k][527].cpu().numpy())
        res = {'workload': '32 wavelengths x 1024 pokes (32x32 Gaussian DM, 1e-3 rad), pupils 256^2/512^2/1024^2 '
                           '-> focal 128^2/256^2/512^2, complex64, f16x2, one poke-set execution',
               'pokes': ps.num_pokes, 'ms': round(sec * 1e3, 3), 'pokes_per_s': round(ps.num_pokes / sec, 1),
               'executed_tflops': round(ps.flops_windowed / sec / 1e12, 1),
               'full_transform_equivalent_tflops': round(ps.flops_full / sec / 1e12, 1),
               'executed_tflop': round(ps.flops_windowed / 1e12, 3), 'full_transform_tflop': round(ps.flops_full / 1e12, 3),
               'device_gb': round(ps.nbytes / 1e9, 1), 'tile_shapes': list(ps.tiles)}
        ps.close()
        del ps, items
        torch.cuda.empty_cache()
        return res

    def time_darkhole_gradient(reps=10):
        """Differentiable dark-hole optimisation step (the broadband Lyot-chain gradient): the 1024^2
        Lyot chain of workloads.lyot_1024 (D = 1 m, f = 1 m, lambda_0 = 500 nm, focal mask radius 3
        lambda_0/D on 256^2, Lyot stop 0.95 D, final focal grid 512^2 at q = 4) at 5 wavelengths over
        a 10 % band; a 48 x 48 Gaussian DM with heights N(0, 1 nm); loss = sum over the wavelengths
        of the dark-zone (3 < r < 10 lambda_0/D, x > 1 lambda_0/D) intensity sum, FP64 reduction.
        One step = DM surface, 5 chains forward, loss, backward to the 2304 actuator heights
        (hand-written adjoint kernels behind torch.autograd), timed eagerly with the host work."""
        lam0 = 500e-9
        wls = lam0 * np.linspace(0.95, 1.05, 5)
        pupil_l, _, _, E_l, fpm, stop, ap = workloads.lyot_1024()
        fpm_grid = ha.make_focal_grid(32, 4, spatial_resolution=lam0)
        focal_l = ha.make_focal_grid(4, 64, spatial_resolution=lam0)
        xf, yf = focal_l.x / lam0, focal_l.y / lam0
        r = np.hypot(xf, yf)
        dz_np = (r > 3) & (r < 10) & (xf > 1)
        dz = torch.as_tensor(dz_np, device=dev)
        count = int(dz_np.sum())
        phase0 = np.angle(E_l) * (ap > 0)
        IFd = ha.make_gaussian_influence_functions(pupil_l, 48, 1.0 / 48)
        heights = np.random.default_rng(5).standard_normal(48 * 48) * 1e-9
        res = {}
        # each precision as the per-wavelength loop (the reference's usage) and with the five
        # wavelengths as grouped launches (LyotCoronagraph / FraunhoferPropagator.forward_multiple)
        scales = torch.as_tensor(4 * np.pi / wls, dtype=torch.float64, device=dev)
        for cdt, precision, key, grouped in ((torch.complex64, 'f16x2', 'f16x2', False),
                                             (torch.complex64, 'f16x2', 'f16x2_grouped_wavelengths', True),
                                             (torch.complex128, None, 'fp64', False),
                                             (torch.complex128, None, 'fp64_grouped_wavelengths', True)):
            fpm_f = ha.Field(torch.as_tensor(fpm, device=dev), fpm_grid)
            stop_f = ha.Field(torch.as_tensor(stop, device=dev), pupil_l)
            coro = ha.LyotCoronagraph(pupil_l, fpm_f, stop_f, focal_plane_mask_grid=fpm_grid, precision=precision)
            prop_l = ha.FraunhoferPropagator(pupil_l, focal_l, precision=precision)
            Es = {wl: torch.as_tensor(ap * np.exp(1j * phase0 * lam0 / wl), device=dev).to(cdt) for wl in wls}
            E_stack = torch.stack([Es[wl] for wl in wls])
            dm = ha.DeformableMirror(IFd)

            heights_dev = torch.as_tensor(heights, device=dev)

            def step():
                act = heights_dev.clone().requires_grad_(True)
                dm.actuators = act
                surface = dm.surface
                total = 0
                if grouped:
                    # the wavelengths as rows of stacked tensors: one phase launch with a scale per
                    # row, grouped transforms, one intensity and one batched dark-zone mean
                    Ep = ha.pointwise.multiply_phase(E_stack, surface, scales)
                    F = prop_l.forward_stacked(coro.forward_stacked(Ep, list(wls)), list(wls))
                    total = ha.get_mean_intensity_in_roi(ha.pointwise.intensity(F), dz).sum() * count
                else:
                    for wl in wls:
                        Ep = ha.pointwise.multiply_phase(ha.Field(Es[wl], pupil_l), surface, 4 * np.pi / wl)
                        img = prop_l.forward(coro.forward(ha.Wavefront(ha.Field(Ep, pupil_l), wl))).intensity
                        total = total + ha.get_mean_intensity_in_roi(img, dz) * count
                total.backward()
                return act.grad

            for _ in range(2):
                step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                g = step()
            torch.cuda.synchronize()
            sec = (time.perf_counter() - t0) / reps
            res[key] = {'ms_per_gradient': round(sec * 1e3, 3), 'gradients_per_s': round(1.0 / sec, 1),
                        'grad_norm': float(g.norm().item())}
            # the same step (forward and backward) captured once in a HIP graph and replayed
            try:
                g_eager = g.clone()
                s = torch.cuda.Stream()
                s.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(s):
                    step()
                torch.cuda.current_stream().wait_stream(s)
                torch.cuda.synchronize()
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph):
                    g_graph = step()
                for _ in range(2):
                    graph.replay()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    graph.replay()
                torch.cuda.synchronize()
                gsec = (time.perf_counter() - t0) / reps
                res[key]['graph_ms_per_gradient'] = round(gsec * 1e3, 3)
                res[key]['graph_gradients_per_s'] = round(1.0 / gsec, 1)
                res[key]['graph_equals_eager'] = bool(torch.equal(g_graph, g_eager))
                del graph
            except Exception as ex:  # reported, not fatal: the eager numbers above stand
                res[key]['graph_error'] = '%s: %s' % (type(ex).__name__, str(ex)[:200])
        res['workload'] = ('1024^2 Lyot chain x 5 wavelengths, 48x48 Gaussian DM, dark-zone loss, gradient to '
                           'the actuators (eager with host work, and graph replays); *_grouped_wavelengths: the '
                           'wavelengths as rows of stacked tensors through grouped launches')
        res['grouped_vs_loop_rel_grad_norm_diff'] = {
            p: abs(res[p + '_grouped_wavelengths']['grad_norm'] - res[p]['grad_norm']) / res[p]['grad_norm']
            for p in ('f16x2', 'fp64')}
        res['f16x2_vs_fp64_rel_grad_norm_diff'] = abs(res['f16x2']['grad_norm'] - res['fp64']['grad_norm']) / \
            res['fp64']['grad_norm']
        return res

    if args.path is not None:
        fn = {'poke_jacobian': time_poke_jacobian, 'mixed_grouped': time_mixed_grouped,
              'broadband_grouped': time_broadband, 'fft_2048': time_fft_2048,
              'darkhole_gradient': time_darkhole_gradient}[args.path]
        print(json.dumps({args.path: fn()}))
        return
    if args.paths_only:
        out = time_other_paths()
        out['mixed_grouped'] = time_mixed_grouped()
        out['poke_jacobian'] = time_poke_jacobian()
        out['darkhole_gradient'] = time_darkhole_gradient()
        print(json.dumps(out))
        return

    log('timing ...')
    main_out = _lib.OUTPUT_INTENSITY
    time_path.isolated = {}
    sec_main, kms_main, out_main = time_path(args.precision, args.steps, args.warmup, main_out)
    kms_iso = time_path.isolated[(args.precision, main_out, True)]
    err_main = err(out_main)
    variants = {}
    outs_for_oracle = {'main': out_main}
    for name, prec, output, parity in [
            ('%s_field_c64' % args.precision, args.precision, _lib.OUTPUT_FIELD, True),
            ('%s_field_c128' % args.precision, args.precision, _lib.OUTPUT_FIELD_C128, True),
            ('%s_intensity_complex_form' % args.precision, args.precision, _lib.OUTPUT_INTENSITY, False),
            ('bf16x3_intensity', 'bf16x3', _lib.OUTPUT_INTENSITY, True),
            ('bf16x3_field_c128', 'bf16x3', _lib.OUTPUT_FIELD_C128, True),
            ('fp32_field_c64', 'fp32', _lib.OUTPUT_FIELD, True),
            ('fp32_intensity', 'fp32', _lib.OUTPUT_INTENSITY, True)]:
        if args.no_variants or (prec == args.precision and output == main_out and parity):
            continue
        sec, kms, o = time_path(prec, max(20, args.steps // 4), 5, output, parity)
        variants[name] = {'value': round(1.0 / sec, 2), 'ms_per_propagation': round(sec * 1e3, 5),
                          'form': forms[(prec, parity)],
                          'peak_rel_intensity_error': err(o), 'kernel_ms': [round(v, 5) for v in kms]}
        outs_for_oracle[name] = o
    fp32_ref = variants.get('fp32_intensity') or variants.get('fp32_field_c64')
    broadband = None if args.no_variants else time_broadband()
    paths = None
    if not args.no_variants:
        paths = time_other_paths()
        paths['mixed_grouped'] = time_mixed_grouped()
        paths['poke_jacobian'] = time_poke_jacobian()
        paths['darkhole_gradient'] = time_darkhole_gradient()

    # Roofline of the dominant kernel (stage 1: T = E . M2, 8 flop per complex MAC of the
    # reference's product, the per-unit count the headline is credited with).
    ny, nx = pupil.shape
    nv, nu = focal.shape
    flops_stage1 = 8.0 * ny * nx * nu
    flops_stage2 = 8.0 * nv * ny * nu
    ms1 = kms_main[1]
    achieved = flops_stage1 / (ms1 * 1e-3) / 1e12
    peak_tf = PEAK_TFLOPS[args.precision]
    main_form = forms.get((args.precision, True), 'complex')
    t0_cfg = tiles.get((args.precision, main_form), (-1, -1))[0]
    if main_form == 'parity':
        # the parity form's own stage 1: four real products of (u' halves) x (2 y' halves) x
        # (x' halves), magnitudes counted as the plan pairs them (the FFT-style focal axis has one
        # unpaired sample, so 257 u' values for 512 samples; the parity split forms the unpaired
        # value's row, so stage 1 runs the 256 paired ones)
        hu, hy, hx = nu // 2, ny // 2, nx // 2
        flops_exec1 = 2.0 * 4 * hu * (2 * hy) * hx
    else:
        flops_exec1 = flops_stage1
    # HBM bytes per stage-1 launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # (copy-calibrated), valid for the tile shape they were measured with
    traffic, traffic_src = None, None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'mft_stage1_hbm_traffic.json')
    if args.precision == 'f16x2' and os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        default_cfg = _lib.TILES_R64X64_W4 if main_form == 'parity' else _lib.TILES_64X64_W8
        cfg_used = t0_cfg if t0_cfg >= 0 else default_cfg
        if tj.get('tile_cfg_stage1') == cfg_used and tj.get('form', 'complex') == main_form:
            traffic = tj['fetch_bytes_per_launch'] + tj['write_bytes_per_launch']
            traffic_src = 'profiles/mft_stage1_hbm_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-scaled)' 
    if main_form == 'parity':
        kname = 'mft_stage_kernel (parity stage 1, X_b . F_ab)'
    elif _lib.TILES_C64X64_W4 <= t0_cfg < _lib.TILES_R64X64_W4:
        kname = 'mft_f16x2_kernel (stage 1, E.M2)'
    else:
        kname = 'mft_stage_kernel (stage 1, E.M2)'
    split_products = {'f16x2': 3, 'bf16x3': 6}.get(args.precision, 1)
    roofline = {'bound': 'mfma', 'kernel': kname, 'form': main_form, 'achieved': round(achieved, 2),
                'peak': peak_tf, 'unit': 'TFLOP/s', 'frac': round(achieved / peak_tf, 4), 'traffic': traffic,
                # flops the launched kernel's own algorithm needs, and the MFMA flops it issues
                # (split products included): the parity form needs about a quarter of the
                # reference's count above
                'form_flop_per_launch': flops_exec1,
                'form_achieved': round(flops_exec1 / (ms1 * 1e-3) / 1e12, 2),
                'form_frac': round(flops_exec1 / (ms1 * 1e-3) / 1e12 / peak_tf, 4),
                'mfma_issued_tflops': round(split_products * flops_exec1 / (ms1 * 1e-3) / 1e12, 2),
                'traffic_unit': 'bytes per launch', 'traffic_source': traffic_src,
                # HIP-event durations inside whole propagations (the achieved figures above), and
                # of each kernel alone over back-to-back launches
                'kernel_ms': {'split': round(kms_main[0], 5), 'stage1': round(kms_main[1], 5),
                              'stage2': round(kms_main[2], 5)},
                'kernel_ms_isolated': {'split': round(kms_iso[0], 5), 'stage1': round(kms_iso[1], 5),
                                       'stage2': round(kms_iso[2], 5)},
                'algorithmic_flop_per_launch': flops_stage1}

    cpu = None
    oracle_err = None
    if not args.no_cpu_baseline:
        log('cpu baseline ...')
        from oracle import hcipy_oracle as o
        from threadpoolctl import threadpool_info
        og_p = o.Grid(pupil.delta, tuple(int(v) for v in pupil.dims), pupil.zero)
        og_f = o.Grid(focal.delta, tuple(int(v) for v in focal.dims), focal.zero)
        # error of every timed output against the oracle's FP64 transform of the same
        # complex64-rounded input (peak-relative intensity)
        I_or = np.abs(o.fraunhofer_forward(E64.astype(np.complex128), og_p, og_f)) ** 2
        oracle_err = {}
        for name, out_t in outs_for_oracle.items():
            a = out_t.cpu().numpy()
            Ia = a.astype(np.float64) if not np.iscomplexobj(a) else np.abs(a.astype(np.complex128)) ** 2
            oracle_err[name] = float(np.abs(Ia - I_or).max() / I_or.max())
        mcpu = o.MftCpu(o.Grid(pupil.delta, tuple(int(v) for v in pupil.dims), pupil.zero),
                        o.Grid(focal.scaled(2 * np.pi).delta, tuple(int(v) for v in focal.dims),
                               focal.scaled(2 * np.pi).zero), 'complex64')
        f = E64.copy()
        t = o.time_it(lambda: mcpu.forward(f), t_max=args.cpu_seconds)
        blas = [i for i in threadpool_info() if i.get('user_api') == 'blas']
        b0 = blas[0] if blas else {}
        cpu = {'value': round(1.0 / t, 3), 'unit': 'propagations/s',
               'cores': int(b0.get('num_threads', len(os.sched_getaffinity(0)))),
               'kind': 'port',
               'sample': 'same 1024^2->512^2 complex64 MFT: scipy cgemm x2 with the reference transpose pattern, '
                         'precomputed matrices, median of 5 runs of >= %.1f s (reference _time_it)' % args.cpu_seconds,
               'blas': '%s %s (%s)' % (b0.get('internal_api', '?'), b0.get('version', '?'),
                                       os.path.basename(b0.get('filepath', '?'))),
               'affinity_cores': len(os.sched_getaffinity(0)),
               'OPENBLAS_NUM_THREADS': os.environ.get('OPENBLAS_NUM_THREADS')}
        if paths is not None and 'lyot_chain_1024_f16x2' in paths:
            # Lyot-chain error against the oracle's FP64 chain, relative to the unocculted peak
            lp, lfg, lfocal, lE, lfpm, lstop, lap = workloads.lyot_1024()
            ogp = o.Grid(lp.delta, tuple(int(v) for v in lp.dims), lp.zero)
            ogm = o.Grid(lfg.delta, tuple(int(v) for v in lfg.dims), lfg.zero)
            ogf = o.Grid(lfocal.delta, tuple(int(v) for v in lfocal.dims), lfocal.zero)
            lE64 = lE.astype(np.complex64).astype(np.complex128)
            I_l = np.abs(o.fraunhofer_forward(o.lyot_forward(lE64, ogp, ogm, lfpm, lstop), ogp, ogf)) ** 2
            I_un = np.abs(o.fraunhofer_forward(lE64 * lstop, ogp, ogf)) ** 2
            for key, It in lyot_outs.items():
                paths[key]['peak_rel_error_vs_oracle_unocculted'] = float(
                    np.abs(It.cpu().numpy().astype(np.float64) - I_l).max() / I_un.max())
            paths['lyot_chain_1024_f16x2']['unocculted_peak_over_coronagraphic_peak'] = float(I_un.max() / I_l.max())
        if poke_sample:
            # one 1024^2 poke of the poke batch against the oracle's poke-and-propagate
            pg, fg, wl, E = poke_sample['item']
            ogp = o.Grid(pg.delta, tuple(int(v) for v in pg.dims), pg.zero)
            ogf = o.Grid(fg.delta, tuple(int(v) for v in fg.dims), fg.zero)
            E64 = E.astype(np.complex64).astype(np.complex128)
            IF = o.gaussian_influence_functions(ogp, poke_sample['n_act'], 1.0 / poke_sample['n_act'],
                                                only=[poke_sample['act']])[0]
            Ib = np.abs(o.fraunhofer_forward(E64, ogp, ogf, wl)) ** 2
            Ip = np.abs(o.fraunhofer_forward(o.dm_forward(E64, poke_sample['h'] * IF, wl), ogp, ogf, wl)) ** 2
            paths['poke_jacobian']['peak_rel_error_vs_oracle_sample'] = float(
                max(np.abs(poke_sample['base'] - Ib).max(), np.abs(poke_sample['poke'] - Ip).max()) / Ib.max())

    value = 1.0 / sec_main
    res = {
        'metric': METRIC,
        'value': round(value, 2),
        'unit': 'propagations/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(sec_main * max(1, args.chunk) * 1e3, 5),
        'propagations_per_step': max(1, args.chunk),
        'ms_per_propagation': round(sec_main * 1e3, 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'f16x2': 'fp16x2', 'bf16x3': 'bf16x3', 'fp32': 'fp32'}[args.precision],
        'data': 'synthetic',
        'config': {'workload': 'fraunhofer_mft_1024_to_512_complex64_focal_intensity', 'pupil': [1024, 1024],
                   'focal': [512, 512], 'q': 4, 'num_airy': 64, 'batch': 1, 'precision': args.precision,
                   'output': 'focal intensity (float32) from the FP64 epilogue value',
                   'step': '%d propagations replayed from one HIP graph' % max(1, args.chunk)
                           if not args.no_graph else '%d eager propagations' % max(1, args.chunk),
                   'tile_shapes': {'%s %s form' % k: list(v) for k, v in tiles.items()} or 'default'},
        'peak_rel_intensity_error': err_main,
        'peak_rel_intensity_error_vs_oracle': oracle_err,
        'speedup_vs_fp32': round(value / fp32_ref['value'], 3) if fp32_ref else None,
        'variants': variants,
        'broadband_grouped': broadband,
        'paths': paths,
        'algorithmic_gflop_per_step': round((flops_stage1 + flops_stage2) / 1e9, 4),
        'roofline': roofline,
        'cpu_baseline': cpu,
    }
    print(json.dumps(res))


if __name__ == '__main__':
    main()
