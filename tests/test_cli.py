"""`heddlemd` front end on the CPU: usage and exit codes
(book/src/reference/cli.md), the config / output-path / init / box checks
of `lint` (src/runner.rs:288-611) and the loader's refusals.  No GPU: lint
without --with-gpu touches no device."""
import os

import pytest

from tests.cli_inputs import BIN, run_cli, write_argon, write_water

USAGE = ("usage: heddlemd run     <config-path>\n"
         "       heddlemd lint    <config-path> [--with-gpu]\n"
         "       heddlemd analyze <analysis-path>\n")

pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="heddlemd binary not built")


@pytest.mark.parametrize("args", [(), ("run",), ("frobnicate", "x.in.toml"), ("lint", "a", "b")])
def test_usage_on_bad_arguments(args):
    r = run_cli(*args)
    assert r.returncode == 1
    assert r.stderr == USAGE


def test_lint_accepts_the_argon_bundle(tmp_path):
    cfg, s = write_argon(str(tmp_path))
    r = run_cli("lint", cfg)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "[heddlemd lint] OK"
    assert "10000 particles" in r.stdout
    assert "not checked (re-run with --with-gpu)" in r.stdout


@pytest.mark.parametrize("solver", ["settle", "shake"])
def test_lint_accepts_the_water_bundle(tmp_path, solver):
    cfg, s = write_water(str(tmp_path), solver=solver)
    r = run_cli("lint", cfg)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "8192 constraint groups" in r.stdout


def test_lint_refuses_existing_outputs(tmp_path):
    cfg, _ = write_argon(str(tmp_path))
    open(os.path.join(str(tmp_path), "argon.out.nve.log"), "w").close()
    r = run_cli("lint", cfg)
    assert r.returncode == 1
    assert "[heddlemd lint] FAIL" in r.stdout
    assert "output file already exists" in r.stderr
    assert r.stderr.startswith("error: ")
    r = run_cli("run", cfg)  # refused before any GPU work
    assert r.returncode == 1 and "already exists" in r.stderr


@pytest.mark.parametrize("edit,msg", [
    (("n_steps = 20", "n_steps = 20\nbogus = 1"), "unknown field `bogus`"),
    (("schema_version = 1", "schema_version = 2"), "schema_version"),
    (('kind = "velocity-verlet"', 'kind = "leapfrog"'), "leapfrog"),
    (("cutoff = 8.5e-10", "cutoff = [8.5e-10"), "TOML parse error at line"),
    (('between = ["Ar", "Ar"]', 'between = ["Ar", "Kr"]'), "unknown particle type"),
])
def test_loader_refusals(tmp_path, edit, msg):
    cfg, _ = write_argon(str(tmp_path))
    text = open(cfg).read()
    assert edit[0] in text
    open(cfg, "w").write(text.replace(edit[0], edit[1], 1))
    r = run_cli("lint", cfg)
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr


def test_box_narrower_than_the_cutoff_is_refused(tmp_path):
    cfg, _ = write_argon(str(tmp_path))
    text = open(cfg).read().replace("cutoff = 8.5e-10\nr_switch = 8.5e-10",
                                    "cutoff = 2.0e-9\nr_switch = 2.0e-9")
    open(cfg, "w").write(text)
    r = run_cli("lint", cfg)
    assert r.returncode == 1
    assert "perpendicular width" in r.stderr


def test_config_name_must_end_in_in_toml(tmp_path):
    cfg, _ = write_argon(str(tmp_path))
    other = os.path.join(str(tmp_path), "argon.toml")
    os.rename(cfg, other)
    r = run_cli("lint", other)
    assert r.returncode == 1 and ".in.toml" in r.stderr


@pytest.mark.parametrize("name", ["lj-10000-argon", "spc-water-256", "chignolin",
                                  "spc-water-8192", "spc-water-65536"])
def test_lint_accepts_the_reference_example_bundles(tmp_path, name):
    from tests.cli_inputs import example_bundle
    cfg = example_bundle(name, str(tmp_path))
    r = run_cli("lint", cfg)
    assert r.returncode == 0, r.stdout + r.stderr
    if name == "spc-water-256":
        assert "512 bonds, 256 angles" in r.stdout
    if name.startswith("spc-water-") and name != "spc-water-256":
        n_mol = int(name.split("-")[-1])
        assert f"{3 * n_mol} particles" in r.stdout and f"{n_mol} constraint groups" in r.stdout
    if name == "chignolin":
        assert "2446 particles" in r.stdout
        assert "171 bonds, 299 angles, 567 dihedrals, 758 constraint groups" in r.stdout
