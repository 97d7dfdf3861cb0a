"""MI355X-native SRD/MPCD time step (gfx950 HIP kernels behind a C ABI)."""
