"""The C++ host layer without a device: YAML parsing, the reference's defaults,
validation errors, time-grid rounding and price scaling (qk_config_load)."""
import math
import os
import shutil
import subprocess

import pytest

import quantkos_amd as qk
from quantkos_amd.engine import QKError, load_config

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture
def in_tmp(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    return tmp_path


def write(tmp, text):
    p = tmp / "cfg.yaml"
    p.write_text(text)
    return str(p)


BASE = """
Market:
  Price: 80.0
  Variance: 0.04
  r: 0.03
  q: 0.01
Options:
  OptionType: European
  OptionRight: Put
  StrikePrice: 100.
Model:
  Heston:
    k: 1.5
    theta: 0.05
    sigma: 0.4
    rho: -0.6
Time:
  T_End: 2.0
  Inp_DT: 0.01
MC:
  N_Paths: 5000
Output:
  out_dir: ""
"""


def test_bench_yaml_loads_with_reference_semantics(in_tmp):
    s = load_config(os.path.join(GOLDEN, "bench_asian_qe.yaml"))
    c, p = s.contract, s.payoff
    assert (c.model, c.scheme, c.precision) == (qk.HESTON, qk.QE, qk.FP64)
    assert c.n_steps == 365 and c.n_paths == 5_000_000 and c.seed == 88471920573105
    assert c.S0 == 1.0 and c.price_scale == 100.0            # Normalize_prices (Config.hpp:710-723)
    assert (c.v0, c.r, c.q, c.T) == (0.05, 0.03, 0.02, 1.0)
    assert (c.kappa, c.theta, c.sigma, c.rho) == (5.0, 0.05, 0.5, -0.8)
    assert (p.type, p.right, p.K) == (qk.ASIAN, qk.CALL, 1.0)
    assert not s.compute_greeks and not s.write_paths


def test_defaults_and_unknown_keys(in_tmp):
    s = load_config(write(in_tmp, BASE + "  not_a_key: 3\n"))
    c = s.contract
    assert c.seed == 184467440737095                          # Config.hpp:403
    assert c.n_steps == 200 and abs(c.T - 2.0) < 1e-15
    assert c.scheme == qk.EULER                               # Numerics block optional, Euler default
    assert math.isclose(s.payoff.K, 100.0 / 80.0) and c.price_scale == 80.0
    assert s.payoff.right == qk.PUT
    # Output block default writes paths to outputs/QuantKos.paths unless the name is empty
    s2 = load_config(write(in_tmp, BASE.replace('Output:\n  out_dir: ""\n', "")))
    assert s2.write_paths == 1 and os.path.isdir(in_tmp / "outputs")


def test_time_grid_rounding(in_tmp):
    """TimeConfig::validate (Config.hpp:347-373): N = ceil(T / Inp_DT), dt = T / N."""
    s = load_config(write(in_tmp, BASE.replace("Inp_DT: 0.01", "Inp_DT: 0.3")))
    assert s.contract.n_steps == 7
    s = load_config(write(in_tmp, BASE.replace("T_End: 2.0", "T_End: 1.0").replace("Inp_DT: 0.01",
                                                                                   "Inp_DT: 0.0027397260273972603")))
    assert s.contract.n_steps == 365


def test_bates_greeks_and_precision(in_tmp):
    s = load_config(os.path.join(GOLDEN, "american_greeks.yaml"))
    assert s.payoff.type == qk.AMERICAN and s.payoff.right == qk.PUT
    assert s.contract.precision == qk.FP32 and s.contract.n_steps == 100
    assert s.compute_greeks == 1
    b = s.bumps
    assert (b.delta_gamma, b.vega_vomma, b.vanna, b.rho, b.theta) == (1, 1, 1, 1, 1)
    assert (b.rel_S, b.abs_vol, b.abs_r) == (0.001, 0.01, 0.0001)
    assert math.isclose(b.abs_T, 1 / 365)
    text = open(os.path.join(GOLDEN, "bates_barrier_paths.yaml")).read()
    s = load_config(write(in_tmp, text))
    c = s.contract
    assert c.model == qk.BATES and (c.lambda_j, c.mu_j, c.sigma_j) == (0.1, -0.05, 0.1)
    assert c.n_steps == 4 and s.write_paths == 1
    assert math.isclose(s.payoff.B, 1.3)


def mc(text, lines):
    """Adds lines to the MC block."""
    return text.replace("  N_Paths: 5000\n", "  N_Paths: 5000\n  " + lines + "\n")


@pytest.mark.parametrize("edit,msg", [
    (lambda t: t.replace("Market:", "Marke:"), "Mandatory block 'Market' missing"),
    (lambda t: t.replace("OptionType: European", "OptionType: Bermudan"), "Invalid value 'Bermudan'"),
    (lambda t: t.replace("  Heston:", "  Bates:\n    k: 1.0\n  Heston:"), "both Heston and Bates"),
    (lambda t: mc(t, "RNG_Seed: 0"), "seed cannot be 0"),
    (lambda t: mc(t, "Batch_Size: 9999"), "batch size must be smaller"),
    (lambda t: mc(t, "Normalize_prices: maybe"), "must be a boolean"),
    (lambda t: t.replace("Price: 80.0", "Price: -1"), "[Market.Price] must be positive."),
    (lambda t: t.replace("rho: -0.6", "rho: -1.5"), "must fall within [-1.0, 1.0]"),
    (lambda t: t.replace("Inp_DT: 0.01", "Inp_DT: 3.0"), "must be smaller that total time"),
    (lambda t: t.replace("  Price: 80.0", "\tPrice: 80.0"), "Invalid YAML syntax"),
    (lambda t: t.replace("N_Paths: 5000", "N_Paths: [1, 2]"), "Invalid YAML syntax"),
    (lambda t: t.replace("N_Paths: 5000", "N_Paths: many"), "bad conversion"),
    (lambda t: mc(t, "compute_vega_et_vomma: true\n  volatility_absolute_bump_size: 0.5"),
     "strictly smaller than sqrt(v0)"),
    (lambda t: mc(t, "compute_theta: true\n  time_absolute_bump_size: 3.0"), "strictly smaller than T_End"),
    (lambda t: mc(t, "Precision: FP16"), "Invalid value 'FP16'"),
    (lambda t: mc(t, "LSM_Basis: Monomial2"), "Invalid value 'Monomial2' for MC.LSM_Basis"),
])
def test_validation_errors(in_tmp, edit, msg):
    with pytest.raises(QKError) as e:
        load_config(write(in_tmp, edit(BASE)))
    assert msg in str(e.value)


def test_missing_file(in_tmp):
    with pytest.raises(QKError, match="File not found"):
        load_config(str(in_tmp / "nope.yaml"))


def test_cli_usage_and_fatal_error(in_tmp):
    cli = os.path.join(os.path.dirname(qk.__file__), "quantkos")
    r = subprocess.run([cli, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage: quantkos" in r.stdout
    r = subprocess.run([cli, str(in_tmp / "missing.yaml")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "FATAL ERROR" in r.stderr and "File not found" in r.stderr


def test_lsm_basis_and_distribution_keys(in_tmp):
    """This build's run-time LSM basis keys (LSRegressionBasis names, ConfigFileEnums.hpp:107-112)
    and the reference's analyze_risk_neutral_payoff_distribution key."""
    s = load_config(write(in_tmp, BASE))
    assert (s.lsm.basis, s.lsm.weighted, s.analyze_distribution) == (qk.LAGUERRE_P03, 0, 0)
    s = load_config(write(in_tmp, mc(BASE, "LSM_Basis: LaguerreP04\n  LSM_Weighted_Basis: true\n"
                                           "  analyze_risk_neutral_payoff_distribution: true")))
    assert (s.lsm.basis, s.lsm.weighted, s.analyze_distribution) == (qk.LAGUERRE_P04, 1, 1)


def test_batch_solver_templates_compile(tmp_path):
    """The drop-in batch types (MSolver, RNGManager, PathsMCBatchMem,
    run_forward_all_mc_batches, LSMEngine) instantiate and link against
    libquantkos as the reference's runner uses them (tests/cpp/batch_solver.cpp)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pkg = os.path.dirname(qk.__file__)
    exe = str(tmp_path / "batch_solver")
    r = subprocess.run(["g++", "-O1", "-std=c++20", "-Wall", "-Werror", "-I" + os.path.join(root, "include"),
                        "-I" + os.path.join(pkg, "host"), os.path.join(root, "tests", "cpp", "batch_solver.cpp"),
                        "-o", exe, "-L" + pkg, "-lquantkos", "-Wl,-rpath," + pkg],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(exe)
