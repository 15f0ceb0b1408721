"""configs/*.toml document the benchmark workloads (C0-C4) and their seeds;
regenerating each from its file must give the particles the benchmarks use."""
from pathlib import Path

import numpy as np
import pytest
import tomli

from gcsim import generators as G

CONFIGS = Path(__file__).resolve().parent.parent / "configs"


def _load(name):
    with open(CONFIGS / name, "rb") as fh:
        return tomli.load(fh)


def _two_disk_from(cfg):
    td = dict(cfg["two_disk"])
    n = td.pop("n_per_galaxy")
    return G.two_disk(n, softening=cfg["physics"]["softening"], G=cfg["physics"]["G"],
                      seed=cfg["simulation"]["seed"], **td)


@pytest.mark.parametrize("name,make", [("c0_two_disk_1000.toml", G.config0),
                                       ("c1_two_disk_65536.toml", G.config1)])
def test_two_disk_configs_match_generators(name, make):
    cfg = _load(name)
    got = _two_disk_from(cfg)
    assert len(got[2]) == cfg["simulation"]["n_particles"]
    for a, b in zip(got, make()):
        assert np.array_equal(a, b)


def test_plummer_config():
    cfg = _load("c2_plummer_1048576.toml")
    n, pl = 4096, cfg["plummer"]
    got = G.plummer(n, pl["total_mass"], pl["scale"], cfg["physics"]["G"], cfg["simulation"]["seed"])
    want = G.plummer(n, seed=1)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert cfg["simulation"]["n_particles"] == 1 << 20


def test_batch_config():
    cfg = _load("c4_batch_256.toml")
    sim, col = cfg["simulation"], cfg["collisions"]
    runs = G.collisions(sim["runs"], sim["seed"], col["n_min"], col["n_max"])
    assert runs == G.collisions(256, seed=2026)
    assert all(col["n_min"] <= r["n"] <= col["n_max"] and r["n"] % 2 == 0 for r in runs)


def test_every_config_parses():
    names = sorted(p.name for p in CONFIGS.glob("*.toml"))
    assert len(names) == 5
    for name in names:
        cfg = _load(name)
        assert {"simulation", "physics"} <= set(cfg)
        assert isinstance(cfg["simulation"]["seed"], int)
