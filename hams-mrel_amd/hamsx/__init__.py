"""hamsx: MI355X (gfx950) frequency-domain panel solver with HAMS-MREL's
interface -- Python host mirror over the C-ABI in include/hamsx.h."""
from ._lib import HamsxError, load  # noqa: F401
from .deck import read_control, read_deck, read_pnl  # noqa: F401
from .mesh import Combined, box, combine, cylinder  # noqa: F401
from .solver import FrequencySolver, wave_properties  # noqa: F401
