"""CPU oracle of the sparse statevector engine (test infrastructure only)."""
