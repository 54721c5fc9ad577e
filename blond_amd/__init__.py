"""blond_amd: MI355X-native engine for BLonD's per-turn tracking path.

Module layout mirrors the reference's import paths so that a BLonD main
script switches by replacing ``blond.`` with ``blond_amd.``:

    blond_amd.input_parameters.ring.Ring
    blond_amd.input_parameters.rf_parameters.RFStation
    blond_amd.beam.beam.Beam, Proton
    blond_amd.beam.distributions.bigaussian
    blond_amd.beam.profile.Profile, CutOptions, OtherSlicesOptions
    blond_amd.impedances.impedance_sources.Resonators
    blond_amd.impedances.impedance.InducedVoltageFreq, TotalInducedVoltage
    blond_amd.trackers.tracker.RingAndRFTracker
    blond_amd.utils.bmath  (the ``bm`` singleton)

plus ``blond_amd.trackers.fused.FusedTracker`` (whole-turn fusion).
All compute on the tracking path runs in libblond_mi355x.so (HIP, gfx950).
"""
__version__ = "0.1.0"
