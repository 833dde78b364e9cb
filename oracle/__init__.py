"""CPU oracle of the heat-bath VCI path (test infrastructure only)."""
