"""Oracle: CPU restatement of the reference's tracking path (test infrastructure only)."""
