"""Host-side logic of the package that needs no GPU."""
import numpy as np
from scipy.signal import windows

from hcipy_amd.coronagraphy import periodic_tukey_window


def test_periodic_tukey_window_matches_scipy():
    for n in (4, 8, 31, 32, 64):
        np.testing.assert_allclose(periodic_tukey_window(n), windows.tukey(n, 1, False), rtol=0, atol=1e-15)
