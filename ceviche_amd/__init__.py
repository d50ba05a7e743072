"""ceviche_amd: MI355X-native (gfx950 HIP) drop-in for ceviche's simulation hot path.

Same public names as ceviche/__init__.py:6-12 for the hot path:
    fdfd_ez, fdfd_hz, fdfd_mf_ez, fdtd, jacobian, primitives (sp_solve, sp_mult,
    spsp_mult), modes (get_modes, insert_mode), utils (measure_fields, my_fft,
    get_spectrum, get_max_power_freq, get_spectral_power).
"""
name = "ceviche_amd"
__version__ = '0.1.0'

from .fdfd import fdfd_ez, fdfd_hz, fdfd_mf_ez  # noqa: E402,F401
from ._fdfd_op import SolverOptions, autotune_instance_workers  # noqa: E402,F401
from .fdtd import fdtd  # noqa: E402,F401
from .jacobians import jacobian  # noqa: E402,F401
from . import primitives  # noqa: E402,F401
from . import modes  # noqa: E402,F401
from . import utils  # noqa: E402,F401
