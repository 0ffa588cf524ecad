"""MI355X (gfx950) implementation of HCIPy's Fourier-propagation path.

Host-side Python mirrors HCIPy's Field / Wavefront / FourierTransform / OpticalElement API;
the arithmetic runs in hand-written HIP kernels behind the C ABI of ``libhcipy_amd.so``
(include/hcipy_amd.h).  Fields are torch tensors on the ROCm device.
"""
from .config import Configuration
from .field import (CartesianGrid, Field, RegularCoords, SeparatedCoords, UnstructuredCoords, default_device,
                    evaluate_supersampled, make_focal_grid, make_pupil_grid, make_uniform_grid, set_default_device,
                    subsample_field)
from .fourier import (ComputationalComplexity, FastFourierTransform, FourierTransform, MatrixFourierTransform, MftGroup,
                      ZoomFastFourierTransform,
                      get_fft_parameters, is_fft_grid, make_fft_grid, make_fourier_transform)
from .optics import (AgnosticOpticalElement, Apodizer, ComplexSurfaceApodizer, OpticalElement, OpticalSystem, PhaseApodizer,
                     SurfaceApodizer, Wavefront)
from .propagation import AngularSpectrumPropagator, FraunhoferPropagator, FresnelPropagator
from .fourier_filter import FourierFilter, NearFieldTransferFunction
from .coronagraphy import LyotCoronagraph, MultiScaleCoronagraph, VectorVortexCoronagraph, VortexCoronagraph
from .polarization import CircularRetarder, JonesMatrixOpticalElement, LinearRetarder, PhaseRetarder
from .spectral_noise import (SpectralNoiseFFT, SpectralNoiseFactoryFFT, SpectralNoiseFactoryMultiscale,
                             SpectralNoiseMultiscale, make_power_law_error)
from .aperture import (make_circular_aperture, make_obstructed_circular_aperture, make_rectangular_aperture, make_spider,
                       make_super_gaussian_aperture)
from .aberration import SurfaceAberration, SurfaceAberrationAtDistance, make_high_pass_power_law_error
from .deformable_mirror import (DeformableMirror, GaussianInfluenceFunctions, PokeSet, TipTiltMirror,
                                XineticsInfluenceFunctions,
                                make_xinetics_influence_functions, make_actuator_positions,
                                make_gaussian_influence_functions)
from .metrics import get_mean_intensity_in_roi, get_mean_raw_contrast, get_strehl_from_focal, get_strehl_from_pupil
from .wavefront_sensing import ModulatedPyramidWavefrontSensorOptics, PyramidWavefrontSensorOptics
from .matrix_inversion import inverse_tikhonov
from . import pointwise

__version__ = '0.1.0'
