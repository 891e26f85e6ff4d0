"""The XPRecon command line: -template, JSON surface and exit codes
(src/main.cpp:21-24,62-137).  CPU only: nothing here reaches the GPU."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xpulse_recon_amd", "bin", "XPRecon")


def run(*args, timeout=60):
    if not os.path.exists(CLI):
        pytest.fail("XPRecon not built (python -m xpulse_recon_amd.build)")
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_template_has_reference_keys_and_defaults():
    r = run("-template")
    assert r.returncode == 0
    t = json.loads(r.stdout)
    assert t["recon"] == {"input": "./", "output": "./recon/", "proj_output": "",
                          "method": "method_1b", "it": 1, "sit": 10, "mlog": True,
                          "weight": 1, "weight_factor": pytest.approx(0.9),
                          "normalize": False, "simulation": False}
    assert t["geometry"] == {"so": 100, "sd": 1000, "angle": 360, "heli_offset": 0,
                             "heli_step": 1, "angle_list": []}
    assert t["detector"]["px"] == pytest.approx(0.1)
    assert t["detector"]["module_number"] == 1
    assert t["detector"]["module_angle_offset"] is None
    assert t["method2"] == {"overlap_fct": 1}
    assert set(t["hartmann"]) == {"devx", "devy", "iterations", "iterations_fct",
                                  "iterations_step", "method"}
    assert set(t["ipr"]) == {"iterations", "iterations_fct", "iterations_step", "energy_kev"}
    assert t["xpr"]["accumulator"] == 64 and t["xpr"]["storage"] == "f32"


def test_template_round_trips_through_the_parser(tmp_path):
    t = run("-template").stdout
    f = tmp_path / "p.json"
    t2 = json.loads(t)
    t2["recon"]["input"] = str(tmp_path / "missing")
    f.write_text(json.dumps(t2))
    # parses (no -2); fails later because the input directory does not exist
    r = run(str(f))
    assert r.returncode == 256 - 3, r.stderr


def test_exit_codes():
    assert run().returncode == 256 - 1                       # no arguments
    r = run("{not json")
    assert r.returncode == 256 - 2 and "JSON error" in r.stderr
    assert run('{"recon": {"it": "three"}}').returncode == 256 - 2
    assert run('{"recon": {"method": "simplex"}}').returncode == 256 - 2
    assert run('{"recon": {"sit": 0}}').returncode == 256 - 2
    assert run('{"xpr": {"accumulator": 16}}').returncode == 256 - 2
    assert run("/nonexistent/params.json").returncode == 256 - 2


def test_hartmann_and_ipr_are_rejected_with_a_message(tmp_path):
    r = run(json.dumps({"recon": {"method": "ipr", "input": str(tmp_path)}}))
    assert r.returncode == 256 - 3
