"""adfdtd — MI355X-native drop-in for the 3-D staggered-grid acoustic FDTD
(`run_fdtd!`) of Nano560/acoustic-disguising.

The time stepping runs in libadfdtd.so (hand-written HIP for gfx950); this
package is the host-side mirror of the reference's Julia API.
"""
from .simulation import (CHANNELS, Cpml, Domain, Transceiver, TransceiverGrid, compute_cpml_coefficients,
                         coords, initial_field, make_cpml, make_domain, make_transceiver, p_rec, p_src,
                         reset_cpml_memory, run_fdtd, t, transceivers_to_grid, trilinear_interpolation, vn_rec,
                         concat_grids,
                         vn_src, x, xv, y, yv, z, zv)
from .geometry import build_scatterer, build_update_mask, outside_predicate
from .illumination import c_for_n, fibonacci_sphere, illumination_points, impulsive_wavelet, r2_sphere, ricker_wavelet
from .extrapolation import PFromPin, PVFromPin, PVFromPV
from .solver import Solver, dtype_code
from .batch import Batch
from .server import Server
from .autotune import tune_batch, tune_solver
from .kh import kh_analytic_accumulate
from .greens import impulsive_gfs, scatterer_params
from .reverb import build_reverb_data, julia_step_range, resample_traces
from .greens_io import (fib_sphere_area, gf_to_physical, gf_to_raw_fdtd, gf_unit_factors, load_gf_mdd, load_gf_ref,
                        read_gf_file, scaling_factors, write_gf_file)
from .hologram import HologramMode, MDDMode, RealMode, RefMode, build_hologram, hologram_mode, slice_indices

__all__ = [n for n in dir() if not n.startswith("_")]
