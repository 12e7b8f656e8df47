"""jht_amd -- MI355X-native HEALPix on-grid synthesis / adjoint (drop-in for jht's pair),
plus the off-grid pair synthesis_general / adjoint_synthesis_general.

The operators are hand-written HIP kernels for gfx950 in ``libjhtx.so`` (C ABI in
``include/jhtx.h``) driven from torch tensors.  Importing this package does not touch
the GPU; the first transform loads the library and builds a plan.
"""

from .healpix import (  # noqa: F401
    RingInfo,
    adjoint_synthesis,
    adjoint_synthesis_batch,
    alm_column_base,
    alm_metric_weight,
    alm_size,
    clear_plans,
    enable_looped_fft,
    get_azimuth_fft_mode,
    get_plan,
    set_azimuth_fft_mode,
    set_tuning,
    synthesis,
    synthesis_batch,
)
from .graph import Instance, RaggedGroup, TransformGraph  # noqa: F401
from .offgrid import (  # noqa: F401
    adjoint_synthesis_general,
    adjoint_synthesis_general_real,
    synthesis_general,
    synthesis_general_real,
)
from .tune import autotune  # noqa: F401

from .callers import (  # noqa: F401
    alm_to_real,
    analysis,
    analysis_real,
    bandpower,
    bare_analysis,
    constrained_realization,
    deconvolve,
    map2alm,
    n_dof,
    pixel_weights,
    pseudo_alm,
    real_to_alm,
    ring_weights,
    synthesis_real,
    wiener,
)

__version__ = "0.1.0"
