"""The Tester drop-in's configuration handling (no suite is run here)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import tester  # noqa: E402


def test_suite_names_are_the_reference_registry():
    assert set(tester.SUITES) == {"RNG", "IOBin", "NoNoise", "BlackScholes", "Heston", "Bates", "AmericanOption",
                                  "EUGreeksPutCallParity"}


def test_selection_rules(tmp_path):
    p = tmp_path / "t.yaml"
    p.write_text("run_tests:\n  - Heston\n  - RNG\n")
    assert tester.select(["x", str(p)]) == ["Heston", "RNG"]
    p.write_text("run_tests:\n  - Heston\n  - Quantum\n")
    assert tester.select(["x", str(p)]) is None
    p.write_text("other: 1\n")
    assert tester.select(["x", str(p)]) is None
    assert tester.select(["x", str(tmp_path / "missing.yaml")]) is None
    assert tester.select(["x"]) == sorted(tester.SUITES)


def test_every_selection_names_an_existing_test():
    for _, nodes in tester.SUITES.values():
        for node in nodes:
            f, _, name = node.partition("::")
            text = open(os.path.join(HERE, f)).read()
            assert not name or ("def " + name + "(") in text, node


def test_cli_help_and_unknown_name(tmp_path):
    r = subprocess.run([sys.executable, os.path.join(HERE, "tester.py"), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "AmericanOption" in r.stdout
    p = tmp_path / "t.yaml"
    p.write_text("run_tests:\n  - Nope\n")
    r = subprocess.run([sys.executable, os.path.join(HERE, "tester.py"), str(p)], capture_output=True, text=True)
    assert r.returncode == 255 and "Unknown test name requested: 'Nope'" in r.stderr
