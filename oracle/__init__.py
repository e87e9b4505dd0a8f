"""CPU oracle for the HighMap grid-erosion hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
bench's cpu_baseline leg, always as the checker.  The product package
(highmap_amd) never imports it.
"""
