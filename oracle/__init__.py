"""CPU oracle of the DBA time step (test infrastructure only; see oracle.h)."""
