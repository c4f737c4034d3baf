"""Host code under AddressSanitizer + UBSan and ThreadSanitizer, no GPU needed.

The library's threaded host code -- the request loop's lanes (csrc/server.cpp), the pipeline's reader, solver
lanes and ordered writer (csrc/pipeline.cpp), the file readers and the mmap writer (csrc/io.cpp,
csrc/mapped_file.hpp), configuration and validation (csrc/config.cpp) and the CLI (cli/main.cpp) -- is compiled
unchanged with g++ against a host stand-in of the device side (tests/sanitize: HIP runtime calls on host memory,
kernel-launching entry points solved by the oracle), then driven by tests/sanitize/host_driver.cpp: several
threads submit, wait and poll the request loop at once, pipelines run with bounded queues and grouped small
cases, a pipeline fails in the middle, lanes fail to initialise, and the readers see truncated, ragged and
malformed files.  Every solve is compared bit for bit with a direct oracle solve.  A sanitizer report fails the
run (halt_on_error); ASan also checks for leaks.
"""
import concurrent.futures as cf
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
CSRC = os.path.join(ROOT, "jacobi-svd-cuda_amd", "csrc")
BUILD = os.path.join(SAN, "_build")
INC = ["-I", SAN, "-I", os.path.join(ROOT, "tests", "simt_model"), "-I", os.path.join(ROOT, "include"), "-I", CSRC]
VARIANTS = {
    "asan": ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"],
    "tsan": ["-fsanitize=thread"],
}
LIBRARY = [os.path.join(CSRC, f) for f in ("config.cpp", "server.cpp", "pipeline.cpp", "io.cpp", "kernels.cpp")]
STUB = os.path.join(SAN, "device_stub.cpp")
ENV = {
    "asan": dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"),
    "tsan": dict(os.environ, TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1"),
}
SCENARIOS = ["server", "server_budget", "server_lane_failure", "pipeline", "pipeline_error", "io"]


def _compile(variant, src):
    out = os.path.join(BUILD, variant, os.path.basename(src) + ".o")
    flags = VARIANTS[variant]
    if src.endswith(".c"):
        cmd = ["gcc", "-std=c11", "-O1", "-g", *flags, "-c", src, "-o", out]
    else:
        cmd = ["g++", "-std=c++20", "-O1", "-g", *flags, "-Wno-unknown-pragmas", "-Wno-psabi", *INC, "-c", src,
               "-o", out]
    subprocess.run(cmd, check=True)
    return out


@pytest.fixture(scope="module")
def binaries():
    srcs = [*LIBRARY, STUB, os.path.join(ROOT, "oracle", "jacobi_oracle.c"), os.path.join(SAN, "host_driver.cpp"),
            os.path.join(ROOT, "jacobi-svd-cuda_amd", "cli", "main.cpp")]
    jobs = [(v, s) for v in VARIANTS for s in srcs if not (v == "tsan" and s.endswith("main.cpp"))]
    for v in VARIANTS:
        os.makedirs(os.path.join(BUILD, v), exist_ok=True)
    with cf.ThreadPoolExecutor(8) as ex:
        objs = dict(zip(jobs, ex.map(lambda j: _compile(*j), jobs)))
    out = {}
    for v in VARIANTS:
        common = [objs[(v, s)] for s in srcs[:-2]]
        for name, main in (("driver", srcs[-2]), ("cli", srcs[-1])):
            if (v, main) not in objs:
                continue
            exe = os.path.join(BUILD, v, name)
            subprocess.run(["g++", *VARIANTS[v], "-o", exe, objs[(v, main)], *common, "-lpthread", "-lm"], check=True)
            out[(v, name)] = exe
    return out


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_host_scenarios(binaries, variant, scenario, tmp_path):
    env = dict(ENV[variant])
    if scenario == "server_lane_failure":
        env["JSVD_STUB_FAIL_BATCH_CREATE"] = "1"
    r = subprocess.run([binaries[(variant, "driver")], scenario, str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0 and f"ok {scenario}" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])


def _cli(binaries, *args, cwd=None):
    r = subprocess.run([binaries[("asan", "cli")], *map(str, args)], capture_output=True, text=True, cwd=cwd,
                       env=ENV["asan"], timeout=300)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-6000:]
    return r


def test_cli_under_asan(binaries, tmp_path):
    assert _cli(binaries, "--help").returncode == 0
    assert _cli(binaries, "--version").returncode == 0
    assert _cli(binaries, "--no-such-option").returncode == 2
    assert _cli(binaries, "-s", "zero", tmp_path / "x.txt").returncode == 2
    assert _cli(binaries, tmp_path / "missing.mat").returncode != 0
    src = tmp_path / "in.txt"
    src.write_text("4 1 0\n1 3 1\n0 1 2\n2 0 1\n\n1 2\n3 4\n5 6\n")
    r = _cli(binaries, src, tmp_path / "out.txt", "--json-report", "-e", "1e-12")
    assert r.returncode == 0, r.stderr[-4000:]
    blocks = [b for b in (tmp_path / "out.txt").read_text().split("\n\n") if b.strip()]
    assert len(blocks) == 6  # U, Sigma, V per case
    # an existing output without --force is refused; with it, a *.mat output of the same cases
    assert _cli(binaries, src, tmp_path / "out.txt").returncode != 0
    r = _cli(binaries, "-i", src, "-o", tmp_path / "out.mat", "--output-format", "mat", "-y", "-q")
    assert r.returncode == 0, r.stderr[-4000:]
    assert (tmp_path / "out.mat").stat().st_size > 0
