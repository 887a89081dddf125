"""CPU oracle for the FDTD time loop: test infrastructure only (see fdtd_oracle.py)."""
