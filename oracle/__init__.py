"""Oracle: CPU restatements of gamdpy's per-timestep path (test infrastructure only)."""
