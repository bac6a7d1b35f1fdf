"""MI355X (gfx950) implementation of the SPRUCE time-step hot path."""
