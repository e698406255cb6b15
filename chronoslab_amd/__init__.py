"""chronoslab_amd: MI355X-native AMG solve phase (V-cycle, AMG-PCG, LOBPCG) for ChronosLab.

Host-side setup lives in `hierarchy`; the HIP solve phase is driven through `amg`.
"""
__version__ = "0.1.0"
