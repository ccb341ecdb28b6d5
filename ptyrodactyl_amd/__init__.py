"""MI355X-native multislice 4D-STEM forward model and gradient.

Drop-in for the multislice path of ptyrodactyl: the public names below keep
the reference's signatures (``src/ptyrodactyl/multislice/simulations.py``,
``src/ptyrodactyl/types/electron_types.py``); the per-position multislice
runs in hand-written gfx950 HIP kernels behind the C-ABI declared in
``include/ptyrodactyl_amd.h``.
"""

from .atom_potentials import bessel_kv, crystal_potential_slices, single_atom_potential
from .checked import checked_cbed_image, checked_make_probe, checked_stem4d_sharded, checked_stem_4d
from .distributions import (TRIVIAL, TRIVIAL_DISTRIBUTION, AxisUpdate, Distribution, ReductionMode,
                            combine_axis_updates, create_axis_update, create_distribution,
                            create_trivial_distribution)
from .form_factors import (KirklandParameters, LobatoParameters, atomic_form_factor, create_kirkland_parameters,
                           create_lobato_parameters, kirkland_form_factor, kirkland_potentials,
                           kirkland_projected_potential, load_form_factor_tables, lobato_bandlimited_peak,
                           lobato_form_factor, lobato_potentials, lobato_projected_potential, projected_atom_potential,
                           set_form_factor_table)
from .producers import coherence_to_distribution, position_jitter_to_distribution
from .reduce import apply_distribution, apply_distributions
from .parallelized import (AtomicSliceData, cbed_amplitude_from_atoms, cbed_image_from_atoms, create_atomic_slice_data,
                           stem4d_sharded)
from .physics import interaction_sigma, phase_interaction_parameter, relativistic_wavelength_ang
from .recon import (OPTIMIZERS, multi_slice_multi_modal, single_slice_multi_modal, single_slice_poscorrected,
                    single_slice_ptychography)
from .simulations import (aberration, annular_detector, bind_cbed_axes, cbed_amplitude, cbed_image,
                          decompose_beam_to_modes, make_probe,
                          probe_modes_to_distribution, propagation_func, shift_beam_fourier, stem_4d, stem_4d_windows,
                          transmission_func)
from .workflows import crystal2stem4d, crystal2stem4d_tiled
from .types import (STEM4D, CalibratedArray, CrystalData, DetectorConfig, EnsembleAxes, LossType, MicroscopeConfig, PotentialSlices,
                    ProbeModes, RuntimeCheckError, create_calibrated_array, create_detector_config,
                    create_crystal_data, create_ensemble_axes, create_microscope_config, create_potential_slices, create_probe_modes,
                    create_stem4d)

__all__ = [
    "aberration", "annular_detector", "cbed_amplitude", "cbed_image", "make_probe", "propagation_func", "shift_beam_fourier",
    "stem_4d", "stem_4d_windows", "transmission_func", "interaction_sigma", "phase_interaction_parameter",
    "relativistic_wavelength_ang", "STEM4D", "CalibratedArray", "DetectorConfig", "EnsembleAxes",
    "MicroscopeConfig", "PotentialSlices", "ProbeModes", "RuntimeCheckError", "create_calibrated_array",
    "create_detector_config", "create_ensemble_axes", "create_microscope_config", "create_potential_slices",
    "create_probe_modes", "create_stem4d", "checked_cbed_image", "checked_make_probe", "checked_stem4d_sharded",
    "checked_stem_4d", "AtomicSliceData", "create_atomic_slice_data", "stem4d_sharded", "LossType", "OPTIMIZERS",
    "single_slice_ptychography", "single_slice_poscorrected", "single_slice_multi_modal", "multi_slice_multi_modal",
    "TRIVIAL", "TRIVIAL_DISTRIBUTION", "AxisUpdate", "Distribution", "ReductionMode", "combine_axis_updates",
    "create_axis_update", "create_distribution", "create_trivial_distribution", "coherence_to_distribution",
    "position_jitter_to_distribution", "apply_distribution", "apply_distributions", "bind_cbed_axes",
    "probe_modes_to_distribution", "decompose_beam_to_modes", "cbed_amplitude_from_atoms", "cbed_image_from_atoms",
    "bessel_kv", "crystal_potential_slices", "single_atom_potential", "CrystalData", "create_crystal_data",
    "KirklandParameters", "LobatoParameters", "atomic_form_factor", "create_kirkland_parameters",
    "create_lobato_parameters", "kirkland_form_factor", "kirkland_potentials", "kirkland_projected_potential",
    "load_form_factor_tables", "lobato_bandlimited_peak", "lobato_form_factor", "lobato_potentials",
    "lobato_projected_potential", "projected_atom_potential", "set_form_factor_table", "crystal2stem4d",
    "crystal2stem4d_tiled",
]
