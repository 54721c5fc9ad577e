from .impedance import (InducedVoltageFreq, InducedVoltageTime, InductiveImpedance,
                        TotalInducedVoltage)
from .impedance_sources import Resonators
