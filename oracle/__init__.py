"""CPU oracle of ComputeQL's operator path (test infrastructure only)."""
