"""Pipe creation (creation::createFromPile, src/creation.cpp:184-230) and its
config.yaml section (profiles/config/config_parsing.cpp:150-188)."""
import numpy as np
import pytest

import akmc
from conftest import make_cfg

YAML = """box:
  size: [8, 8, 8]
  lattice_const: 2.85532
  cutoff_radius: 5.6
create:
  random:
    enable: {rnd}
    va_count: 3
    alloy:
      Fe: 9666
      Cu: 134
      Mn: 130
      Ni: 70
  pipe:
    enable: {pipe}
    input_box: {path}
seeds:
  create_types: 100565
  create_vacancy: 16650
  event_selection: 563264
  time_inc: 594689
simulation:
  temperature: 573
  physics_time: 1e-3
  attempt_freq: 6e12
"""


def _atoms(cfg):
    c = dict(cfg)
    c["va_count"] = 0
    return akmc.random_sites(c)


def test_pipe_sites(tmp_path):
    cfg = make_cfg(n=8, va=5)
    base = _atoms(cfg)
    X, Y = 16, 8
    cu = int(np.flatnonzero(base == akmc.BIT_CU)[0])
    cx, cy, cz = cu % X, (cu // X) % Y, cu // (X * Y)
    f = tmp_path / "vac.txt"
    # one site on a Cu atom, one plain, a repeat, two outside the box, then a
    # triple that does not parse: reading stops there
    f.write_text("%d %d %d\n3 4 5\n3 4 5\n-1 0 0\n16 0 0\n0 0 8\nx 1 1\n7 7 7\n" % (cx, cy, cz))
    s = akmc.pipe_sites(cfg, f)
    want = base.copy()
    want[cu] = akmc.BIT_V
    want[3 + X * (4 + Y * 5)] = akmc.BIT_V
    assert np.array_equal(s, want)
    assert int((s & akmc.BIT_V).sum()) == 2
    # the random vacancies of the config play no part
    assert np.array_equal(akmc.pipe_sites(make_cfg(n=8, va=40), f), want)


def test_pipe_file_must_exist(tmp_path):
    with pytest.raises(akmc.AkmcError):
        akmc.pipe_sites(make_cfg(n=8), tmp_path / "missing.txt")


def test_pipe_config(tmp_path):
    y = tmp_path / "config.yaml"
    y.write_text(YAML.format(rnd="false", pipe="true", path="vac.txt"))
    p = akmc.params_from_yaml(str(y))
    assert p.va_count == 0 and p.n_alloy == 4 and list(p.box) == [8, 8, 8]
    assert akmc.pipe_path_from_yaml(y) == "vac.txt"
    y.write_text(YAML.format(rnd="true", pipe="true", path="vac.txt"))   # random wins
    assert akmc.params_from_yaml(str(y)).va_count == 3
    with pytest.raises(akmc.AkmcError):
        akmc.pipe_path_from_yaml(y)
    y.write_text(YAML.format(rnd="false", pipe="false", path="vac.txt"))
    with pytest.raises(akmc.AkmcError):
        akmc.params_from_yaml(str(y))
