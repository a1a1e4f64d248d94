"""MI355X-native SBE propagation path (HIP kernels behind a C ABI)."""
