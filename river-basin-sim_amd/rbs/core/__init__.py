"""The solver core: grid, state, schemes (LI, HLLC), ledger -- reference solver/core."""
