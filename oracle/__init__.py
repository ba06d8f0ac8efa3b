"""FP64 CPU oracle for the per-stage right-hand side (test infrastructure
only: tests/, the smoke check and the benchmark's CPU-baseline leg)."""
from .oracle import Oracle, build, DEFAULT_SCHEME, UP7  # noqa: F401
