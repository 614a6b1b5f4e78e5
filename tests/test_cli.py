"""``python -m gpufsm`` commands (upstream `tests/test_cli.py`): listing and environment on
any machine, verification on the GPU."""

from __future__ import annotations

import pytest

from gpufsm import cli
from gpufsm.core.registry import Backend, is_available


def test_list_names_every_backend_and_hip_techniques(capsys):
    assert cli.main(["list"]) == 0
    out = capsys.readouterr().out
    for b in Backend:
        assert out.count(b.value) >= 1
    assert "hip" in out and "chain" in out and "split" in out


def test_env_reports_library_and_backends(capsys):
    assert cli.main(["env"]) == 0
    out = capsys.readouterr().out
    assert "hip lib" in out and "ABI" in out and "backends" in out


def test_verify_without_a_device_says_why(capsys):
    if is_available(Backend.HIP):
        pytest.skip("a HIP device is visible; the GPU suite runs verify")
    assert cli.main(["verify", "--quick"]) == 1
    assert "no GPU backend" in capsys.readouterr().out


def test_bench_rejects_unknown_example(capsys):
    assert cli.main(["bench", "--example", "nope"]) == 2


@pytest.mark.gpu
def test_verify_on_the_gpu(capsys):
    assert cli.main(["verify"]) == 0
    out = capsys.readouterr().out
    assert "0 failure(s)" in out and "DIFFERS" not in out


@pytest.mark.gpu
def test_bench_example_on_the_gpu(capsys):
    assert cli.main(["bench", "--repeats", "3", "--warmup", "1"]) == 0
    assert "hip/" in capsys.readouterr().out
