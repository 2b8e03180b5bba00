"""MI355X-native Siddon projector / backprojector and MLEM / OSEM update for APIRL.

The compute path is the gfx950 HIP library apirl_amd/libapirl_hip.so (C-ABI in
include/apirl_hip.h); `apirl_amd.hip` binds it (plans, the Projector entry points, the
device-resident EM driver, batches), `apirl_amd.cli` runs the reference's .par-driven
commands on it."""
__all__ = ["hip", "phantoms"]
