"""Configuration flags of the Fourier path (layout of hcipy/config/default_config.yaml:1-59).

``Configuration()`` returns the process-wide singleton; attributes nest like the reference's
(``Configuration().fourier.mft.precompute_matrices``).  Setting an unknown key raises
TypeError (hcipy/config/config.py:21-55).  Additions of this package:

* ``fourier.mft.precision_complex64``: arithmetic of complex64 MFTs, ``'f16x2'`` (split FP16
  MFMA with FP64 flushes, the default), ``'bf16x3'`` (three-way split BF16 MFMA) or ``'fp32'``
  (plain FP32 MFMA).  complex128 always runs in FP64.
* ``fourier.zfft.precision_complex64``: arithmetic of complex64 zoom FFTs, ``'fp64'`` (FP32
  loads and stores around FP64 chirps, line FFTs and intermediate; the default) or ``'fp32'``.
* ``execution_time_prediction_coefficients`` (fft, mft, zfft): the reference's values, so that
  ``make_fourier_transform(planner='estimate')`` makes the reference's choice;
  ``planner='measure'`` times the candidates on the device instead.
"""
import copy

_DEFAULTS = {
    'fourier': {
        'fft': {
            'emulate_fftshifts': True,
            'execution_time_prediction_coefficients': {'a': 3.746, 'b': 0.881, 'c': -4.749},
        },
        'mft': {
            'precompute_matrices': True,
            'allocate_intermediate': True,
            'precision_complex64': 'f16x2',
            'execution_time_prediction_coefficients': {'a': 1.667, 'b': 0.716, 'c': -4.624},
        },
        'zfft': {
            'precision_complex64': 'fp64',
            'execution_time_prediction_coefficients': {'a': 4.093, 'b': 0.847, 'c': -2.973},
        },
    },
    'core': {'use_new_style_fields': False},
}


class _Node:
    def __init__(self, d):
        object.__setattr__(self, '_d', d)

    def __getattr__(self, name):
        try:
            v = self._d[name]
        except KeyError:
            raise AttributeError(name)
        return _Node(v) if isinstance(v, dict) and not _is_leaf_dict(name) else v

    def __setattr__(self, name, value):
        if name not in self._d:
            raise TypeError('Unknown configuration key %s' % name)
        self._d[name] = value

    def __getitem__(self, name):
        return self._d[name]


def _is_leaf_dict(name):
    return name == 'execution_time_prediction_coefficients'


class Configuration(_Node):
    _instance_data = None

    def __init__(self):
        if Configuration._instance_data is None:
            Configuration._instance_data = copy.deepcopy(_DEFAULTS)
        object.__setattr__(self, '_d', Configuration._instance_data)

    def reset(self):
        Configuration._instance_data.clear()
        Configuration._instance_data.update(copy.deepcopy(_DEFAULTS))
