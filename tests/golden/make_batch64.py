"""Writes batch64_instances.json: the batched-serving workload's 64 instances
as drawn by tcscf_amd.configs.batch_instances() from its fixed seed.  The
fixture pins the generator (tests/test_abi_cpu.py re-draws and compares)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tcscf_amd import configs  # noqa: E402

if __name__ == "__main__":
    with open(os.path.join(HERE, "batch64_instances.json"), "w") as f:
        json.dump(configs.batch_instances(), f, indent=1)
        f.write("\n")
