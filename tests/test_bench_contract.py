"""bench.py's command-line contract on the CPU: one GPU only, workload configurations."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_workload_configurations():
    import bench
    sc, p, cfg = bench.workload("cubes_8")
    assert cfg["workload"] == "cubes_8_on_ground" and sc.n_bodies == 8 and p.ground_enabled
    sc, p, cfg = bench.workload("ref_chain_144", "mixed")
    assert cfg["workload"] == "ref_chain_144" and sc.n_bodies == 144 and p.precision == 1
    assert p.dhat == 0.095 and p.kappa == 1e5 and not p.ccd_enabled  # the reference's own parameters
    assert bench.REF_PUBLISHED["ref_chain_144"] == 0.10
