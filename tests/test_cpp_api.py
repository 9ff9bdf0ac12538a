"""The C++ host interface (include/akmc_sim.hpp: parseConfig, EventHooks,
ABVIModel, simulation::simulate) driven by tests/cpp/sim_api_test.cpp, built
by the package Makefile with g++ against libakmc.so."""
import math
import os
import subprocess

import numpy as np
import pytest

import akmc
import refcpu
from conftest import make_cfg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "misa-akmc-gpu_amd", "lib", "sim_api_test")


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "misa-akmc-gpu_amd"), "lib/sim_api_test"], check=True)
    return BIN


def _yaml(cfg, pipe=None):
    al = "\n".join("      %s: %d" % (k, v) for k, v in cfg["alloy"])
    s = cfg["seeds"]
    return ("box:\n  size: [%d, %d, %d]\n  lattice_const: 2.85532\n  cutoff_radius: 5.6\n"
            "create:\n  random:\n    enable: %s\n    va_count: %d\n    alloy:\n%s\n"
            "  pipe:\n    enable: %s\n    input_box: %s\n"
            "seeds:\n  create_types: %d\n  create_vacancy: %d\n  event_selection: %d\n  time_inc: %d\n"
            "simulation:\n  temperature: %r\n  physics_time: %r\n  attempt_freq: %r\n  sector_time: %r\n"
            "output:\n  dump:\n    interval: 100\n    file_path: \"kmc.{}.out\"\n") % (
        cfg["box"][0], cfg["box"][1], cfg["box"][2], "false" if pipe else "true", cfg["va_count"], al,
        "true" if pipe else "false", pipe or "\"\"", s["create_types"], s["create_vacancy"],
        s.get("event_selection", 0), s["time_inc"], cfg["temperature"], cfg["physics_time"],
        cfg.get("attempt_freq", 6e12), cfg["sector_T"])


def _run(*args, cwd=None):
    print("cpp api:", " ".join(str(a) for a in args), flush=True)
    r = subprocess.run([_binary(), *[str(a) for a in args]], capture_output=True, text=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)


def test_config(tmp_path):
    cfg = make_cfg(n=12, va=7, T=650.0, sector_T=2e-7, physics=3e-6)
    y = tmp_path / "config.yaml"
    y.write_text(_yaml(cfg))
    out = _run("config", y)
    assert out["box"] == "12,12,12" and out["n_alloy"] == "4" and out["va_count"] == "7"
    assert float(out["temperature"]) == 650.0 and float(out["sector_time"]) == 2e-7
    assert float(out["physics_time"]) == 3e-6 and out["create_pipe"] == "0" and out["dump_interval"] == "100"
    y.write_text(_yaml(cfg, pipe="vac.txt"))
    out = _run("config", y)
    assert out["create_pipe"] == "1" and out["pipe_input_box"] == "vac.txt" and out["va_count"] == "0"


def _oracle_final(cfg, sites=None):
    r = refcpu.RefCpu(cfg, create=sites is None)
    if sites is not None:
        r.set_sites(sites)
    r.run(10 ** 9)
    return r.sites(), r.state()


@pytest.mark.gpu
@pytest.mark.parametrize("hook_steps", [1, 3])
def test_simulate_random_creation(tmp_path, hook_steps):
    T = 3e-7
    cfg = make_cfg(n=16, va=30, T=600.0, sector_T=T, physics=7.5 * T)
    y = tmp_path / "config.yaml"
    y.write_text(_yaml(cfg))
    out = _run("simulate", y, hook_steps, tmp_path / "sites.bin")
    steps = math.ceil(7.5)
    assert out["step"] == str(steps) and out["all_done"] == "1" and out["seed_check"] == "ok"
    assert out["hook_steps"] == ",".join(str(s) for s in range(steps) if s % hook_steps == 0)
    assert out["defects"] == "30" and out["dev_vacancy_check"] == "ok"
    sites = np.fromfile(tmp_path / "sites.bin", dtype=np.uint8)
    ref_sites, rs = _oracle_final(cfg)
    assert int(out["events"]) == rs["events"] and rs["step"] == steps
    assert np.array_equal(sites, ref_sites)


@pytest.mark.gpu
def test_simulate_pipe_creation(tmp_path):
    T = 3e-7
    cfg = make_cfg(n=16, va=0, T=620.0, sector_T=T, physics=4 * T)
    rng = np.random.default_rng(4)
    (tmp_path / "vac.txt").write_text(
        "".join("%d %d %d\n" % (rng.integers(32), rng.integers(16), rng.integers(16)) for _ in range(25)))
    y = tmp_path / "config.yaml"
    y.write_text(_yaml(cfg, pipe="vac.txt"))
    out = _run("simulate", y, 1, tmp_path / "sites.bin", cwd=tmp_path)   # the pipe path is relative, as in the reference
    start = akmc.pipe_sites(cfg, tmp_path / "vac.txt")
    assert out["defects"] == str(int((start & akmc.BIT_V).sum()))
    sites = np.fromfile(tmp_path / "sites.bin", dtype=np.uint8)
    ref_sites, rs = _oracle_final(cfg, start)
    assert int(out["events"]) == rs["events"] and out["hook_steps"] == "0,1,2,3"
    assert np.array_equal(sites, ref_sites)


@pytest.mark.gpu
def test_dumbbell_rates_through_the_cpp_layer():
    out = _run("itl")
    assert out["itl_dumbbells"] == "6" and out["itl_check"] == "ok"
