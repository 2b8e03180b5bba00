"""MI355X-native Siddon projector / backprojector and MLEM / OSEM update for APIRL.

The compute path is the gfx950 HIP library apirl_amd/libapirl_hip.so (C-ABI in
include/apirl_hip.h); `apirl_amd.hip` binds it, `apirl_amd.recon` mirrors the
reference's Projector / Mlem / Osem interface on top of it."""
__all__ = ["hip", "phantoms"]
