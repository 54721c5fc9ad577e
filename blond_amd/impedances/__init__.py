from .impedance import InducedVoltageFreq, TotalInducedVoltage
from .impedance_sources import Resonators
