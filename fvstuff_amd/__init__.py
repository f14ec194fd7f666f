"""MI355X-native 2-D compressible-Euler DGSEM step (HIP kernels behind a C ABI)."""
