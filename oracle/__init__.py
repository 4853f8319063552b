"""ORACLE package — test infrastructure only (CPU restatement of rustracer)."""
