"""CPU checks of the boundary: the C-ABI library loads without a GPU, exports
every symbol include/aspher_hip.h declares, and the Python mirror keeps the
reference's signatures and defaults (python/bindings.cpp:804-879)."""
import inspect
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "aspher_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(aspher_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from aspher_amd import _capi
    L = _capi.lib()
    decl = _declared()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(_capi.EXPORTS) == decl


def test_library_loads_and_reports_without_gpu():
    from aspher_amd import _capi
    assert _capi.lib().aspher_version().startswith(b"aspher")


def test_contact_solver_signature_matches_reference():
    import aspher
    sig = inspect.signature(aspher.ContactSolver.__init__)
    want = dict(domain_size=1.0, E_star=1.0, eta=2.0, aca_tol=1e-6, leaf_size=64,
                use_hmatrix=True, use_acagp=False, central_fraction=0.3, inline_svd_tol=0.0,
                backend="", q=4, near_radius=1, h2_leaf_side=8)
    for k, v in want.items():
        assert sig.parameters[k].default == v, k
    ssig = inspect.signature(aspher.ContactSolver.solve)
    for k, v in dict(tol=1e-8, max_iter=5000, use_pr=True, precond="none", p_init=None).items():
        assert ssig.parameters[k].default == v, k


def test_solve_nested_signature_matches_reference():
    import aspher
    sig = inspect.signature(aspher.solve_nested)
    want = dict(domain_size=1.0, E_star=1.0, coarsest=64, q=6, leaf_side=8, precond=True,
                precond_engine="auto", precond_radius=2, precond_occupancy_max=0.4,
                precond_cost_gate=0.0, tol=1e-8, coarse_tol=1e-4, max_iter=20000, use_pr=True,
                single_precision=False, precision="", float_floor=2e-6,
                allow_tolerance_relaxation=False, light_result=False, backend="h2",
                record_error_history=False, active_set=False, active_delta=0.05,
                active_halo=2, active_max_rounds=5, active_all_levels=True,
                active_occupancy_max=0.4)
    names = list(sig.parameters)
    assert names[:3] == ["grid_size", "gap", "p_nominal"]
    for k, v in want.items():
        assert sig.parameters[k].default == v, k


def test_constructor_validation_precedes_device_use():
    import aspher
    for kw in (dict(grid_size=0), dict(grid_size=8, domain_size=0.0),
               dict(grid_size=8, E_star=-1.0), dict(grid_size=8, backend="nope"),
               dict(grid_size=8, backend="h2"), dict(grid_size=8)):
        with pytest.raises(ValueError):
            aspher.ContactSolver(**kw)


def test_hmatrix_contact_alias():
    import hmatrix_contact
    import aspher
    assert hmatrix_contact is aspher


def test_product_package_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "aspher_amd")
    pat = re.compile(r"^\s*(import|from)\s+oracle\b|liboracle|oracle/", re.M)
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert not pat.search(txt), os.path.join(dirpath, f)


def test_ledger_record_keys(tmp_path):
    """Ledger records carry the reference harness's keys (bench/harness.py:514-560)."""
    import json
    from types import SimpleNamespace
    import numpy as np
    from aspher_amd import ledger
    res = SimpleNamespace(
        contact_area=0.1, memory=dict(cg_state=1, best_iterate=2, contact_mask=3,
                                      output_arrays=4, peak=10),
        timings=dict(total=1.0, matvec=0.5, precond=0.1, build=0.2, coarse=0.1,
                     verification=0.0, candidate=0.0, output=0.0),
        status="converged", status_reason="", converged=True, iterations=10,
        matvec_count=21, verification_matvec_count=0, precond_count=10,
        precond_dropped=False, identification_steps=0, requested_tol=1e-6,
        effective_tol=1e-6, validation_scope="solve_operator", operator_error_kind="unavailable",
        fw_error=1e-7, penetration_error=1e-7, load_error=0.0, pk_error=1e-6,
        mean_pressure=0.005, active_rounds=0, active_fallback=False,
        stage_stats=[dict(name="finest:64")], pressure=np.ones(16))
    rec = ledger.solve_record("w", 4, "v", dict(p_bar=0.005), np.zeros(16), dict(grid_size=4),
                              res, [1.0, 0.5, 0.7])
    expect = {"workload", "Ns", "N", "variant", "surface_sha", "workload_params", "p_bar",
              "occupancy", "solver_args", "wall_cold_s", "wall_warm_s", "wall_all_s",
              "surface_gen_s", "solver_memory", "timings", "status", "status_reason",
              "converged", "iterations", "matvec_count", "verification_matvec_count",
              "precond_count", "precond_dropped", "identification_steps", "requested_tol",
              "effective_tol", "validation_scope", "operator_error_kind", "fw_error",
              "penetration_error", "load_error", "pk_error", "contact_area", "mean_pressure",
              "active_rounds", "active_fallback", "stage_stats", "pressure_sha", "run_status",
              "timestamp"}
    assert set(rec) == expect
    assert rec["wall_warm_s"] == 0.7 and rec["pressure_sha"] == ledger.digest(np.ones(16))
    path = tmp_path / "l.jsonl"
    ledger.append(str(path), rec, prov={"revision": "x"})
    back = json.loads(path.read_text().strip())
    assert back["provenance"] == {"revision": "x"} and back["N"] == 16
