"""MI355X-native Debye-SAXS engine: hand-written gfx950 HIP kernels behind a C ABI.

The compute path lives in libresaxs_amd.so (resaxs_amd/csrc); this package is
its Python handle (ctypes) plus seeded workload generators.
"""
from ._lib import (RsxError, RSX_DOTS_DODECA_ICOSA, RSX_DOTS_FIBONACCI, RSX_VARIANT_FP32,  # noqa: F401
                   RSX_VARIANT_FP32R, RSX_VARIANT_SHARED, RSX_VARIANT_WAVE)
from .engine import (Engine, atomic_form_factors, atomic_radii, density_dots, device_count,  # noqa: F401
                     device_name, dodeca_icosa_dots, fibonacci_dots, fit_ensemble, make_bodies, pdb_atom_type,
                     read_pdb, sphere_dots)
