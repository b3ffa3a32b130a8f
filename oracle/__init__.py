"""Test infrastructure: CPU restatement of the reference hider (see refhider.h)."""
