"""CPU oracle of the HCI hot path -- test infrastructure only (see oracle.cpp)."""
