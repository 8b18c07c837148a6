"""MI355X-native label-indexed anisotropic elastic FDTD."""
