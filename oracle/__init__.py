"""Oracle package: CPU restatement of the reference solver (test infrastructure only)."""
