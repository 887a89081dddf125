"""fdtdx_amd: an MI355X-native FDTD time-loop engine behind the fdtdx host API.

The public names mirror the reference package (ymahlau/fdtdx 0.6.2, src/fdtdx/__init__.py) for
the objects, configuration and run functions the time loop needs.
"""

from . import constants
from .config import GradientConfig, SimulationConfig, UniformGrid
from .core import GaussianWindow, OnOffSwitch, TemporalWindow, TukeyWindow, WaveCharacter
from .detectors import (ClosedSurfacePhasorPoyntingFluxDetector, ClosedSurfacePoyntingFluxDetector,
                        CustomModeOverlapDetector, EnergyDetector, FieldDetector, GaussianModeOverlapDetector,
                        ModeOverlapDetector, PhasorDetector, PhasorPoyntingFluxDetector, PoyntingFluxDetector, gaussian_mode_fields,
                        gaussian_mode_function, normalize_by_poynting_flux)
from .fdtd import (ArrayContainer, FieldState, ObjectContainer, ParameterContainer, apply_params, backward,
                   checkpointed_fdtd, full_backward,
                   custom_fdtd_forward, place_objects, reversible_fdtd, run_fdtd)
from .materials import DispersionModel, DrudePole, LorentzPole, Material
from .objects import (BlochBoundary, BoundaryConfig, Device, GridCoordinateConstraint, PerfectElectricConductor,
                      PerfectlyMatchedLayer, PerfectMagneticConductor, PositionConstraint, SimulationVolume,
                      SizeConstraint, SizeExtensionConstraint, UniformMaterialObject, boundary_objects_from_config)
from .recorder import DtypeConversion, LinearReconstructEveryK, Recorder
from .modes import compute_mode
from .sources import (GaussianPlaneSource, GaussianPulseProfile, ModePlaneSource, PointDipoleSource,
                      SingleFrequencyProfile, UniformPlaneSource)

from .symmetry import field_component_parity, unfold_detector_states, unfold_fields
__all__ = [n for n in dir() if not n.startswith("_")]
