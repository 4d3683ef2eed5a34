"""CPU oracle of the KAPSEL SPM hot path -- test infrastructure only."""
