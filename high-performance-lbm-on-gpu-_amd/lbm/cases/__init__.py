"""Validation cases with exact or literature answers, run on the HIP backends
(src/lbm/cases/ of the reference)."""
