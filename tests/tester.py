"""Drop-in for the reference's Tester binary (tests/Tester.cpp, TestEngine.hpp:60-170):
runs the named suites listed under `run_tests:` in a YAML file (all suites when no
file is given), refuses unknown names, and returns the number of failed suites
(-1 on a configuration error).  Each suite is the set of this repository's tests
that restate the reference suite of that name; they run in one pytest process.

    python tests/tester.py [TestserConfig.yaml] [--help]
"""
from __future__ import annotations

import os
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))

# Suite name (tests/Tester.cpp:76-98) -> description, pytest node selections.
SUITES = {
    "RNG": ("Verify that standard gaussian random variable is sampled correctly",
            ["test_oracle_rng.py", "test_gpu_parity.py::test_gpu_normal_moments"]),
    "IOBin": ("Check that data are saved to disk and reloaded correctly using Binary File Format.",
              ["test_gpu_host.py::test_paths_file_matches_device_paths"]),
    "NoNoise": ("Zero-Variance SDE Drift",
                ["test_gpu_parity.py::test_no_noise_drift", "test_gpu_host.py::test_no_noise_read_back"]),
    "BlackScholes": ("Weak Convergence to Black Scholes", ["test_gpu_parity.py::test_black_scholes_limit"]),
    "Heston": ("Weak Convergence to Heston",
               ["test_gpu_parity.py::test_heston_closed_form", "test_gpu_parity.py::test_config0_contract"]),
    "Bates": ("Weak Convergence to Bates", ["test_gpu_parity.py::test_bates_closed_form"]),
    "AmericanOption": ("American Option LSM", ["test_gpu_parity.py::test_american_ikonen_toivanen"]),
    "EUGreeksPutCallParity": ("Check that EU option Greeks obey put-call parity",
                              ["test_gpu_parity.py::test_greek_put_call_parity"]),
}
INDENT = "   "


def help_message() -> str:
    lines = ["Usage: python tests/tester.py [YAML_CONFIG_FILE] [--help]", "",
             "To run specific tests, create a YAML file with a 'run_tests' sequence:", "",
             "run_tests:", '  - "test_id_1"', '  - "test_id_2"', "",
             "If no YAML file is provided, the engine will execute all registered tests.",
             "List Available tests :"]
    lines += ["%s - %s  #  %s" % (INDENT, k, v[0]) for k, v in sorted(SUITES.items())]
    return "\n".join(lines)


def select(argv):
    """The suite names to run, or None after a configuration error (printed)."""
    if len(argv) > 1:
        try:
            with open(argv[1]) as f:
                cfg = yaml.safe_load(f) or {}
        except (OSError, yaml.YAMLError):
            print("Error: Could not parse YAML file or file not found.\n" + help_message(), file=sys.stderr)
            return None
        names = cfg.get("run_tests")
        if not names:
            print("[ FATAL ERROR ] Test configuration is empty, 'run_tests' node not found!\n" + help_message(),
                  file=sys.stderr)
            return None
        names = [str(n) for n in names]
    else:
        names = sorted(SUITES)
    for n in names:
        if n not in SUITES:
            print("%s[ FATAL ERROR ] Unknown test name requested: '%s'\n%sAborting test execution to prevent a "
                  "silent pass." % (INDENT, n, INDENT), file=sys.stderr)
            return None
    return names


class _Outcomes:
    def __init__(self):
        self.failed, self.ran = set(), set()

    def pytest_runtest_logreport(self, report):
        if report.when == "call" or report.failed:
            self.ran.add(report.nodeid)
            if report.failed:
                self.failed.add(report.nodeid)


def run(argv) -> int:
    if "--help" in argv[1:]:
        print(help_message())
        return 0
    names = select(argv)
    if names is None:
        return -1
    import pytest
    sel = sorted({node for n in names for node in SUITES[n][1]})
    out = _Outcomes()
    pytest.main(["-q", "-p", "no:cacheprovider", "-o", "addopts="] + [os.path.join(HERE, s) for s in sel],
                plugins=[out])
    n_failed = 0
    for n in names:
        mine = [nid for nid in out.ran if any(nid.startswith("tests/" + s) or ("/" + s) in nid or nid.startswith(s)
                                             for s in SUITES[n][1])]
        bad = [nid for nid in mine if nid in out.failed]
        if mine and not bad:
            print("%s[ SUCCESS ] %s Suite passed successfully.\n" % (INDENT, SUITES[n][0]))
        else:
            print("%s[ TEST FAILURE ] %s Suite detected an error!\n" % (INDENT, SUITES[n][0]))
            n_failed += 1
    print("%s STATUS   : %s" % (INDENT, "ALL PASSED OK" if n_failed == 0 else
                                "FAILED (%d suite(s) broke constraints)" % n_failed))
    return n_failed


if __name__ == "__main__":
    sys.exit(run(sys.argv))
