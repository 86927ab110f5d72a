"""The workloads the fixtures cover (defined once in mptc_amd/workloads.py)."""
from mptc_amd.workloads import BATCH_STREAMS, SEQUENCES, SynthCfg, synth_cfg  # noqa: F401
