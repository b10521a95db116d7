"""MI355X-native Lanczos engine for diabatic vibronic Hamiltonians in a
harmonic-oscillator direct-product basis (drop-in for the Lanczos path of
NadVibS).  The compute path is libnadvibs.so (HIP, gfx950); this package holds
its Python binding and the reference-compatible host side."""
from .inputs import Problem, read_problem, write_problem  # noqa: F401
