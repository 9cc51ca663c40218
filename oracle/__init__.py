"""CPU oracle for the prism path: test infrastructure only (never imported by jif3d_amd)."""
