"""Host-side numerical helpers of the SH hot path (geometry, binning operators, FITS, inputs)."""
