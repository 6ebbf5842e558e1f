"""MI355X-native SPH + self-gravity giant-impact step (HIP/gfx950 behind a C ABI)."""
