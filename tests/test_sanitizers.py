"""Race and memory checks of the host code (CPU).

* The oracle's threaded step -- the reference's pthread scheme of worker
  threads, barriers and serial sections elected by the barrier
  (src/main.c:747-839, libmpm/process_usl.c:465-592) -- run on four threads
  over small jobs that reach every threaded path (oracle/race_check.c: USL
  and MUSL, g_local_mu2, g_nonlocal, isolin, dp_ri_mt, box and square_silo
  walls, the per-crossing log lines, particles leaving the grid), built with
  ThreadSanitizer and with AddressSanitizer + UndefinedBehaviorSanitizer.
* The driver's configuration reader (mpm-2d_amd/host/cfg.c, through
  mpm_usl_run --check-cfg) on files in the reference's libconfuse syntax
  and on malformed ones, built with AddressSanitizer +
  UndefinedBehaviorSanitizer (leak checking on).

A sanitizer report ends the program with a non-zero status.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
ORACLE_BUILD = os.path.join(REPO, "oracle", "_build")
ASAN_DRIVER = os.path.join(REPO, "mpm-2d_amd", "bin", "asan", "mpm_usl_run")

CASES = ("usl_box_g_local_mu2", "musl_box_g_local_mu2", "usl_silo_falling",
         "usl_box_g_nonlocal", "usl_box_isolin", "usl_box_dp_ri_mt")


def _env(**kw):
    env = dict(os.environ)
    env.update(kw)
    return env


def _run_race_check(binary, tmp_path, env):
    path = os.path.join(ORACLE_BUILD, binary)
    assert os.path.exists(path), f"{path} is not built (run __graft_entry__.build())"
    r = subprocess.run([path, str(tmp_path / "crossings.log")], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
    lines = dict(line.split(": ", 1) for line in r.stdout.strip().split("\n"))
    assert set(lines) == set(CASES)
    stats = {}
    for name, text in lines.items():
        words = text.replace(",", "").split()
        stats[name] = dict(np=int(words[0]), crossings=int(words[4]), active=int(words[6]))
    # the paths meant to be reached were reached: cells crossed (log lines
    # written), particles deactivated after leaving through the orifice
    assert stats["usl_box_g_local_mu2"]["crossings"] > 0
    assert stats["musl_box_g_local_mu2"]["crossings"] > 0
    falling = stats["usl_silo_falling"]
    assert 0 < falling["active"] < falling["np"]
    assert os.path.getsize(tmp_path / "crossings.log") > 0
    return stats


def test_oracle_threaded_step_is_race_free(tmp_path):
    _run_race_check("race_check_tsan", tmp_path,
                    _env(TSAN_OPTIONS="halt_on_error=1 exitcode=66 second_deadlock_stack=1"))


def test_oracle_threaded_step_is_memory_clean(tmp_path):
    _run_race_check("race_check_asan", tmp_path,
                    _env(ASAN_OPTIONS="detect_leaks=1 halt_on_error=1 exitcode=66",
                         UBSAN_OPTIONS="halt_on_error=1 print_stacktrace=1"))


CFG_GOOD = """timestep
{
    dt = 1e-05   // comment after a value
    automatic-dt = 0
}
/* block
   comment */
solver { solver-type = explicit-usl }
material
{
    material-file = "g_local_mu2.so"
    use-builtin = 0
    properties = {1e6, 0.3, 0.38}
    integer-properties = { }
}
boundary-conditions
{
    boundary-conditions-file = "box.so"
    use-builtin = 0
    properties = {0, 0, 1, 1}
    integer-properties = {1, 3, 1, 0}
}
input { initial-particle-file = "p.txt" grid-file = "g.txt" }
output
{
    directory = "out"
    user = ${MPM_TEST_UNSET_VAR:-unknown}
    sample-rate = 120
}
"""

CFG_BAD = {
    "unterminated_list.cfg": ("material { properties = {1, 2 \n", 0x20),
    "builtin_material.cfg": ("material { use-builtin = 1 }\n", 0x21),
    "unknown_solver.cfg": ("solver { solver-type = explicit-xyz }\nmaterial { use-builtin = 0 }\n"
                           "boundary-conditions { use-builtin = 0 }\n", 0x20),
    "unterminated_string.cfg": ('material { material-file = "abc\n', 0x20),
    "unbalanced_braces.cfg": ("timestep { dt = 1e-5 \n", 0x20),
    "stray_brace.cfg": ("}\n", 0x20),
    "unterminated_comment.cfg": ("timestep { dt = 1e-5 } /* no end", 0x20),
    "unterminated_env.cfg": ("output { user = ${NO_END\n}\n", 0x20),
    "missing_include.cfg": ('include("does-not-exist.cfg")\n', 0x20),
    "self_include.cfg": ('include("self_include.cfg")\n', 0x20),
    "bad_include.cfg": ("include(\n", 0x20),
    "missing_value.cfg": ("timestep { dt = }\n", 0x20),
}


def _asan_env():
    return _env(ASAN_OPTIONS="detect_leaks=1 halt_on_error=1 exitcode=66",
                UBSAN_OPTIONS="halt_on_error=1 print_stacktrace=1")


def test_cfg_reader_is_memory_clean(tmp_path):
    assert os.path.exists(ASAN_DRIVER), f"{ASAN_DRIVER} is not built"
    good = tmp_path / "simulation.cfg"
    good.write_text(CFG_GOOD)
    r = subprocess.run([ASAN_DRIVER, "-c", str(good), "--check-cfg", "0.25"],
                       capture_output=True, text=True, timeout=120, env=_asan_env())
    assert r.returncode == 0, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
    got = dict(line.split("=", 1) for line in r.stdout.strip().split("\n"))
    assert float(got["dt"]) == 1e-5 and got["automatic-dt"] == "0"
    assert got["solver"] == "explicit-usl" and got["user"] == "unknown"
    assert [int(v) for v in got["bc-integer-properties"].split(",")] == [1, 3, 1, 0]
    assert float(got["sample-rate"]) == 120.0 and float(got["t_stop"]) == 0.25


@pytest.mark.parametrize("name", sorted(CFG_BAD))
def test_cfg_reader_errors_are_memory_clean(tmp_path, name):
    text, code = CFG_BAD[name]
    path = tmp_path / name
    path.write_text(text)
    r = subprocess.run([ASAN_DRIVER, "-c", str(path), "--check-cfg"], capture_output=True,
                       text=True, timeout=120, env=_asan_env(), cwd=tmp_path)
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == code, (r.returncode, r.stderr[-2000:])
