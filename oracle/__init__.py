"""FP64 CPU oracle of the pllcart3d time step -- test infrastructure only."""
