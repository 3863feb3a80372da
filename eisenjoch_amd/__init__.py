"""MI355X-native electrostatic global placer (the ElectroPlace hot path of
Eisenjoch), hand-written HIP kernels for gfx950 behind a C ABI."""
from .placer import (  # noqa: F401
    AdmissionRefused,
    Context,
    Device,
    DeviceError,
    ElectroPlaceCfg,
    FlatDesign,
    InitialPlacementFailed,
    NoBelsAvailable,
    NonFiniteError,
    PlacementFailed,
    PlaceResult,
    PlacerElectro,
    PlacerError,
    Server,
    batch_designs,
    batch_sizes,
    bins_for_side,
    config1,
    config2,
    config3,
    generate,
    legalize,
    place_electro,
    sized_design,
    timing_net_weights,
    total_hpwl_int,
    with_spacers,
)

__version__ = "0.1.0"
