"""bench.py's command line on the CPU: more than one GPU (by flag or by a
torchrun world size) is refused with exit code 2 and a message, before
anything touches a GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("argv,env", [(["--gpus", "2"], {}), (["--gpus", "8"], {}),
                                      ([], {"WORLD_SIZE": "2"})])
def test_bench_refuses_several_gpus(argv, env):
    e = dict(os.environ, **env)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *argv],
                       capture_output=True, text=True, env=e, timeout=60)
    assert p.returncode == 2
    assert "measures one GPU only" in p.stderr
    assert p.stdout == ""
