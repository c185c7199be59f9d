"""CPU oracle package (test infrastructure only; never imported by vecgeom_amd)."""
