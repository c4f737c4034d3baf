"""Command-line front end (jacobi-svd-cuda_amd/cli/main.cpp) on the host: the reference CLI's option table,
output naming, argument errors and exit codes (reference src/main.cpp:877-1300).  Nothing here touches the GPU
(help, version, dry runs and argument errors return before any solve)."""
import os
import subprocess

import pytest

CLI = os.path.join(os.path.dirname(__file__), "..", "jacobi-svd-cuda_amd", "lib", "jacobi-svd")


def run(*args, cwd=None):
    if not os.path.exists(CLI):
        pytest.fail("CLI not built (python jacobi-svd-cuda_amd/build.py)")
    p = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=60)
    return p.returncode, p.stdout, p.stderr


def dry(*args):
    rc, out, err = run(*args, "--dry-run")
    assert rc == 0, err
    lines = dict(ln.split(":", 1) for ln in out.splitlines()[1:] if ":" in ln)
    return {k.strip(): v.strip() for k, v in lines.items()}


def test_help_and_version():
    rc, out, _ = run("--help")
    assert rc == 0 and "Usage:" in out
    for name in ("--input PATH", "--output PATH", "--input-format FMT", "--output-format FMT", "--format FMT",
                 "--epsilon NUM", "--max-sweeps N", "--threads-per-block N", "--layout-transpose-mode MODE",
                 "--layout-transpose-min-cols N", "--layout-transpose-min-elems N", "--layout-transpose-auto-tune",
                 "--layout-transpose-bench-reps N", "--layout-transpose-bench-sweeps N", "--queue-capacity N",
                 "--force", "--dry-run", "--print-config", "--json-report", "--quiet", "--help", "--version"):
        assert name in out, name
    for short in ("-i, --input", "-o, --output", "-f, --format", "-e, --epsilon", "-s, --max-sweeps",
                  "-t, --threads-per-block", "-c, --queue-capacity", "-y, --force", "-q, --quiet", "-h, --help",
                  "-v, --version"):
        assert short in out, short
    rc, out, _ = run("-v")
    assert rc == 0 and out.startswith("jacobi-svd-cuda CLI v0.1.0")


def test_defaults_and_output_naming(tmp_path):
    a = tmp_path / "case.mat"
    a.write_bytes(b"")
    cfg = dry(a)
    assert cfg["output"] == str(tmp_path / "case.svd.mat")
    assert (cfg["input-format"], cfg["output-format"]) == ("auto", "mat")
    assert (cfg["epsilon"], cfg["max-sweeps"], cfg["threads-per-blk"]) == ("1e-09", "128", "256")
    assert (cfg["layout-mode"], cfg["layout-min-cols"], cfg["layout-min-elem"]) == ("auto", "16", "4096")
    assert (cfg["layout-auto-tune"], cfg["layout-bench-rep"], cfg["layout-bench-swp"]) == ("false", "2", "8")
    assert (cfg["queue-capacity"], cfg["force-overwrite"]) == ("4", "false")
    t = tmp_path / "in.txt"
    t.write_text("")
    assert dry(t)["output"] == str(tmp_path / "in.svd.txt")
    # an output without extension takes the input's format and its extension
    assert dry(t, tmp_path / "res")["output"] == str(tmp_path / "res.txt")
    assert dry("--format", "mat", tmp_path / "in.txt")["output"] == str(tmp_path / "in.svd.mat")
    cfg = dry("-i", a, "-o", tmp_path / "o.mat", "-e1e-12", "-s", "7", "-c", "2", "--layout-transpose-mode=off",
              "--layout-transpose-auto-tune", "-yq")
    assert (cfg["epsilon"], cfg["max-sweeps"], cfg["queue-capacity"]) == ("1e-12", "7", "2")
    assert (cfg["layout-mode"], cfg["layout-auto-tune"], cfg["force-overwrite"]) == ("off", "true", "true")
    assert dry(a, "--solver", "mixed")["solver"] == "mixed"
    assert dry(a)["device-batch"] == "0" and dry(a, "--device-batch", "64")["device-batch"] == "64"


@pytest.mark.parametrize("args,message", [
    ((), "Missing input file. Use --input <path>."),
    (("{d}/none.mat",), "Input file does not exist"),
    (("{d}/x.bin",), "Cannot infer input format from extension. Use --input-format {{mat|txt}}."),
    (("{d}/a.mat", "{d}/a.mat"), "Input and output paths must be different files."),
    (("{d}/a.mat", "{d}/exists.mat"), "Output file already exists. Use --force to overwrite"),
    (("{d}/a.mat", "--epsilon", "0"), "Option --epsilon must be a positive finite number."),
    (("{d}/a.mat", "--epsilon", "abc"), "Invalid floating-point value for --epsilon."),
    (("{d}/a.mat", "--max-sweeps", "7x"), "Invalid integer value for --max-sweeps."),
    (("{d}/a.mat", "--max-sweeps", "-3"), "Option --max-sweeps must be positive."),
    (("{d}/a.mat", "--queue-capacity", "0"), "Option --queue-capacity must be positive."),
    (("{d}/a.mat", "--format", "csv"), "Invalid value for --format: expected one of {{auto, mat, txt}}."),
    (("{d}/a.mat", "--layout-transpose-mode", "maybe"), "expected one of {{auto, on, off}}"),
    (("{d}/a.mat", "--bogus"), "Unknown option: --bogus"),
    (("{d}/a.mat", "-Z"), "Unknown short option: -Z"),
    (("{d}/a.mat", "--force=yes"), "Option --force does not take a value."),
    (("{d}/a.mat", "--epsilon"), "Option --epsilon requires a value."),
    (("{d}/a.mat", "{d}/b.mat", "{d}/c.mat"), "Unexpected positional argument"),
    (("{d}/a.mat", "{d}/b.txt", "--solver", "mixed"), "--solver block|mixed reads and writes *.mat files only."),
    (("{d}/a.mat", "--device-batch", "0"), "Option --device-batch must be positive."),
    (("{d}/a.mat", "--solver", "block", "--device-batch", "8"), "--device-batch applies to the plain solver only."),
])
def test_argument_errors_exit_2(tmp_path, args, message):
    (tmp_path / "a.mat").write_bytes(b"")
    (tmp_path / "x.bin").write_bytes(b"")
    (tmp_path / "exists.mat").write_bytes(b"")
    rc, out, err = run(*[x.format(d=tmp_path) for x in args])
    assert rc == 2
    assert err.startswith("Argument error: ") and message.format() in err, err
    assert "Use --help for usage." in err


def test_execution_failure_exit_1(tmp_path):
    # a truncated *.mat header is an I/O error of the pipeline, reported with exit code 1 (no GPU is reached)
    a = tmp_path / "bad.mat"
    a.write_bytes(b"\x00\x00\x00")
    rc, out, err = run(a, "-q")
    assert rc == 1 and err.startswith("Execution failed: "), err
