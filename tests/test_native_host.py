"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU suite, no GPU).

Builds checks with g++/gcc ``-fsanitize=address,undefined`` (and one with
``-fsanitize=thread``) and runs them:

* ``tests/native/host_check.cpp`` with ``csrc/layout.cpp`` (the device-layout builders,
  which parse untrusted CSR input) and ``oracle/fsm_oracle.c`` (the C oracle): the epsilon
  fold, rows, gather lists, DFA tables and chain tiles checked against an independent set
  simulation and the oracle on random automata, plus malformed-input refusals;
* ``tests/native/capi_check.cpp`` with ``csrc/capi.hip`` compiled as host C++ and the
  library's kernel objects (linked, never launched): the C ABI's host-only paths;
* ``tests/native/pool_check.cpp`` with ``csrc/copy_pool.hpp`` under ThreadSanitizer: the
  serving scheduler's staging copy pool.

Any sanitizer report aborts the program, which fails the test.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "abstraction-regret_amd" / "csrc"
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
       "-fno-omit-frame-pointer"]
# host C++ view of the HIP sources: the HIP headers' host declarations, plus the two macros
# the HIP compiler itself provides (__align__ as the GNU attribute, HIP_SYMBOL as the
# symbol's address)
HIP_HOST = ["-D__HIP_PLATFORM_AMD__", "-D__align__(x)=__attribute__((aligned(x)))",
            "-DHIP_SYMBOL(x)=(&(x))", "-I/opt/rocm/include", f"-I{CSRC}"]
KERNEL_OBJECTS = ["dfa.o", "nfa.o", "chain.o", "general.o"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="g++/gcc not installed")


def _run(cmd, **kw):
    res = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert res.returncode == 0, f"{' '.join(map(str, cmd))}\n{res.stdout}\n{res.stderr}"
    return res


@pytest.fixture(scope="module")
def objs(tmp_path_factory):
    out = tmp_path_factory.mktemp("native")
    _run(["g++", "-std=c++17", *SAN, *HIP_HOST, "-c", str(CSRC / "layout.cpp"),
          "-o", str(out / "layout.o")])
    _run(["gcc", "-std=c11", *SAN, "-pthread", "-c", str(ROOT / "oracle" / "fsm_oracle.c"),
          "-o", str(out / "fsm_oracle.o")])
    return out


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


def test_layouts_and_oracle_under_sanitizers(objs):
    exe = objs / "host_check"
    _run(["g++", "-std=c++17", *SAN, *HIP_HOST, "-c", str(ROOT / "tests/native/host_check.cpp"),
          "-o", str(objs / "host_check.o")])
    _run(["g++", "-fsanitize=address,undefined", str(objs / "host_check.o"),
          str(objs / "layout.o"), str(objs / "fsm_oracle.o"), "-pthread", "-o", str(exe)])
    res = _run([str(exe)], env=_env(), timeout=300)
    assert "host_check ok" in res.stdout


def _kernel_objects() -> list[str]:
    build = CSRC / "_build"
    if not all((build / o).exists() for o in KERNEL_OBJECTS):
        _run([sys.executable, str(CSRC / "build.py")], timeout=900)
    return [str(build / o) for o in KERNEL_OBJECTS]


def test_capi_host_paths_under_sanitizers(objs):
    if not Path("/opt/rocm/lib/libamdhip64.so").exists():
        pytest.skip("the HIP runtime library is not installed")
    kernels = _kernel_objects()
    exe = objs / "capi_check"
    _run(["g++", "-std=c++17", *SAN, *HIP_HOST, "-x", "c++", "-c", str(CSRC / "capi.hip"),
          "-o", str(objs / "capi.o")])
    _run(["g++", "-std=c++17", *SAN, *HIP_HOST, "-c", str(ROOT / "tests/native/capi_check.cpp"),
          "-o", str(objs / "capi_check.o")])
    _run(["g++", "-fsanitize=address,undefined", str(objs / "capi_check.o"),
          str(objs / "capi.o"), str(objs / "layout.o"), *kernels, "-L/opt/rocm/lib",
          "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-pthread", "-o", str(exe)])
    res = _run([str(exe)], env=_env(), timeout=300)
    assert "capi_check ok" in res.stdout


def test_copy_pool_under_thread_sanitizer(tmp_path):
    """The serving scheduler's staging copy pool (``csrc/copy_pool.hpp``) under
    ThreadSanitizer: successive copies into one guarded buffer, pools of 0 to 7 helpers."""
    exe = tmp_path / "pool_check"
    _run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", f"-I{CSRC}",
          str(ROOT / "tests/native/pool_check.cpp"), "-pthread", "-o", str(exe)])
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    res = _run([str(exe)], env=env, timeout=300)
    assert "pool_check ok" in res.stdout
