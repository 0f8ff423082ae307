"""Oracle: CPU restatement of the reference's algorithm (test infrastructure only)."""
