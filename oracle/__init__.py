"""CPU oracle of the particle-tracing path (test infrastructure only)."""
