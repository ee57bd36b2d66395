"""MI355X (gfx950) implementation of the UMWM timestep hot path."""
