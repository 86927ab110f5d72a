"""mptc_amd: MI355X-native MPTC per-frame encode path (DXT1 fit, index-block
reuse search, endpoint re-fit) behind a C-ABI (include/mptc.h)."""
