from .beam_feedback import BeamFeedback
