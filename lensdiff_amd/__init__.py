"""MI355X-native LensDiff diffraction backend (HIP, gfx950)."""
