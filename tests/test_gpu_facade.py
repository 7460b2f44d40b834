"""The C++ facades driven as the reference drives them: sixteen
bbfmm::kernel objects per operator application combined as in
electric.cpp:294-388, and GMRES(std::function) (gmres.cpp:50-164) on that
operator, against the plan's matvec and device GMRES
(tests/native/facade_check.cpp, built into ibim-levelset_amd/bin)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ibim-levelset_amd", "bin", "facade_check")


def test_facades_match_the_plan(gpu):
    out = subprocess.run([EXE, "2000", "5"], capture_output=True, text=True, timeout=240)
    print(out.stdout[-2000:])
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-2000:]
    assert "FACADE CHECK OK" in out.stdout
    assert "CONVERGED" in out.stdout  # the reference's GMRES banner lines
