"""Host side of the drop-in executable without a GPU: options file
(Config.cpp:15-49), PQR reading (molecule.h:21-51), the reference's scaled grid
(main.cpp:33-45) and the Angstrom grid, via `levelset --dry-run`."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ibim-levelset_amd", "bin", "levelset")

PQR = """REMARK   1 two atoms
ATOM      1  N   ALA     1      -1.000   0.500   2.000 -0.4000 1.8500
ATOM      2  CA  ALA     1       1.500  -0.250   2.000  0.6000 1.7000
HETATM    3  O   HOH     2       9.000   9.000   9.000  0.0000 1.4000
"""


def dry(tmp_path, lines):
    (tmp_path / "m.pqr").write_text(PQR)
    cfg = tmp_path / "in.cfg"
    cfg.write_text("\n".join(lines) + "\n")
    out = subprocess.run([EXE, "--dry-run", str(cfg)], capture_output=True, text=True, timeout=30, cwd=tmp_path)
    return out


def test_config_and_pqr_reference_grid(tmp_path):
    out = dry(tmp_path, ["# a comment", "pqr_file = m.pqr", "grid_size = 128", "broken line", "", "fmm_np = 5"])
    assert out.returncode == 0, out.stderr
    assert out.stdout.count("Incomplete configuration") == 1
    assert re.search(r"fmm_np = 5", out.stdout) and re.search(r"pqr_file = m.pqr", out.stdout)
    assert re.search(r"ATOMS\s+2", out.stdout)          # HETATM lines are not atoms
    assert re.search(r"NET CHARGE\s+0.2", out.stdout)
    assert re.search(r"GRID 128 x 128 x 128 from -300 -300 -300", out.stdout)
    # half extent: x spans -1-1.85 .. 1.5+1.7, i.e. 3.025; h = (600/128) / (240/3.025)
    h = float(re.search(r"h\s+(\S+) Angstroms", out.stdout).group(1))
    assert abs(h - (600 / 128) / (240 / 3.025)) < 1e-5


def test_angstrom_grid_and_synthetic(tmp_path):
    out = dry(tmp_path, ["synthetic_atoms = 500", "synthetic_seed = 500", "synthetic_radius = 20", "grid_h = 0.25"])
    assert out.returncode == 0, out.stderr
    assert re.search(r"ATOMS\s+500", out.stdout)
    q = float(re.search(r"NET CHARGE\s+(\S+)", out.stdout).group(1))
    assert abs(q + 2.0) < 1e-9
    assert re.search(r"h\s+0.25 Angstroms", out.stdout)


def test_missing_molecule_is_an_error(tmp_path):
    out = dry(tmp_path, ["pqr_file = nothere.pqr"])
    assert out.returncode != 0 and "pqr_file" in out.stderr
