"""TEST INFRASTRUCTURE ONLY: FP64 CPU oracle of the UMWM hot path."""
