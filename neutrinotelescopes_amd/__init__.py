"""neutrinotelescopes_amd — MI355X-native photon propagation for neutrino telescopes.

A drop-in for the photon-propagation path of the Julia package
NeutrinoTelescopes: the same source, medium, spectrum, target and hit-record
types, propagated by a hand-written HIP kernel for gfx950 (libntprop.so,
C ABI in include/ntprop.h).
"""
from .medium import (WaterProperties, MediumProperties, make_cascadia_medium_properties,
                     refractive_index, group_velocity, dispersion, scattering_length,
                     absorption_length, cherenkov_angle, DIPPR105)
from .spectrum import (Monochromatic, CherenkovSpectrum, InvLambda2Spectrum, frank_tamm,
                       frank_tamm_norm, frank_tamm_inverted_cdf, integrate_gauss_quad)
from .lightyield import (Particle, PEPlus, PEMinus, PGamma, LongitudinalParameters,
                         LongitudinalParameterisation, PointlikeIsotropicEmitter,
                         PointlikeTimeRangeEmitter, AxiconeEmitter, PencilEmitter,
                         ExtendedCherenkovEmitter, PointlikeCherenkovEmitter, split_source,
                         oversample_source, total_lightyield, cherenkov_track_length)
from .detection import (DetectionSphere, MultiPMTDetector, make_pom_pmt_coordinates, lambertian_angular_acceptance,
                        check_pmt_hit, area_acceptance, p_one_pmt_acc, get_pmt_count)
from .propagation import (PhotonPropSetup, PhotonHits, PropagationEngine, propagate_photons, resample_simulation,
                          run_photon_prop_no_local_cache, calc_total_weight, calc_time_residual,
                          make_hits_from_photons, get_engine)

from . import refflow  # noqa: E402  (the reference flow's input layout and model file)
from . import pmt_frontend  # noqa: E402  (PMT front end: waveforms and unfolding on the device)
from .pmt_frontend import PMTConfig, PMTFrontEnd, make_reco_pulses, std_pmt_config  # noqa: E402

__all__ = [n for n in dir() if not n.startswith("_")]
