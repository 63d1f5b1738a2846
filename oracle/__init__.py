"""CPU oracle for dsDecompX's Host communication path (test infrastructure only)."""
