"""gamma_amd: MI355X-native (HIP/gfx950) implementation of GAMMA's per-step
hydro hot path behind a C-ABI (include/gammahip.h)."""
