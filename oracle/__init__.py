"""Host restatement of the reference rasteriser: the checker (test infrastructure only)."""
