"""MI355X (gfx950) CKKS hot path: negacyclic NTT over 64-bit RNS towers,
ModUp/ModDown basis conversion and Hybrid key switching, behind the C ABI of
include/ntt_hip.h."""
from . import hip, params  # noqa: F401
from .hip import Context, DeviceBuffer, KeySwitch, Event, Stream, ContractError, NttHipError  # noqa: F401
