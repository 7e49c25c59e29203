"""JSONL benchmark ledger records in the reference's format.

The reference's harness (bench/harness.py:214-247 provenance, :474-561
worker record) appends one JSON object per timed solve to a ledger, with the
provenance of the binary and machine, the workload and a hash of its
surface, every solver argument, the result's status, tolerances, errors,
phase timings, byte accounting and stage statistics, and a hash of the
returned pressure. `solve_record` emits the same keys, so a record of this
library can be read next to the reference's records by the same analysis
tools.
"""
from __future__ import annotations

import hashlib
import json
import os
import platform
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _run(cmd) -> str:
    try:
        return subprocess.run(cmd, capture_output=True, text=True, timeout=10).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return ""


def digest(a) -> str:
    """blake2b-128 of a float64 array's bytes (the reference's surface_sha /
    pressure_sha)."""
    v = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return hashlib.blake2b(v.view(np.uint8), digest_size=16).hexdigest()


def provenance() -> dict:
    """Machine and build provenance (bench/harness.py:214-247)."""
    from . import build
    from ._capi import lib
    dirty = _run(["git", "-C", ROOT, "status", "--porcelain", "--untracked-files=no"])
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    mem = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    mem = round(int(line.split()[1]) * 1024 / 2 ** 30, 2)
    except OSError:
        pass
    hipcc = build._hipcc()
    ver = _run([hipcc, "--version"]).splitlines()
    return {
        "revision": _run(["git", "-C", ROOT, "rev-parse", "--short", "HEAD"]),
        "dirty": bool(dirty),
        "dirty_files": len(dirty.splitlines()) if dirty else 0,
        "cxx": hipcc,
        "cxx_version": ver[0] if ver else "",
        "cxx_flags": " ".join(f for f in build.HIPCC_FLAGS if not f.startswith("-I")),
        "build_type": "Release",
        "fft_engine": "aspher-rb (register-blocked radix-16, HIP gfx950)",
        "library": lib().aspher_version().decode(),
        "offload_arch": build.ARCH,
        "cpu": cpu,
        "cores_online": os.cpu_count(),
        "affinity": sorted(os.sched_getaffinity(0)),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS", ""),
        "mem_available_gib": mem,
        "python": platform.python_version(),
        "host": platform.node(),
    }


def solve_record(workload: str, Ns: int, variant: str, workload_params: dict, gap,
                 solver_args: dict, res, times, surface_gen_s: float = 0.0) -> dict:
    """One ledger record of a timed solve (bench/harness.py:514-560 keys)."""
    warm = sorted(times[1:])
    return {
        "workload": workload, "Ns": Ns, "N": Ns * Ns, "variant": variant,
        "surface_sha": digest(gap), "workload_params": workload_params,
        "p_bar": workload_params.get("p_bar"),
        "occupancy": float(res.contact_area),
        "solver_args": {k: v for k, v in solver_args.items() if k not in ("gap", "gap_ptr")},
        "wall_cold_s": times[0],
        "wall_warm_s": warm[len(warm) // 2] if warm else None,
        "wall_all_s": list(times),
        "surface_gen_s": surface_gen_s,
        "solver_memory": dict(res.memory),
        "timings": dict(res.timings),
        "status": res.status, "status_reason": res.status_reason,
        "converged": bool(res.converged),
        "iterations": int(res.iterations),
        "matvec_count": int(res.matvec_count),
        "verification_matvec_count": int(res.verification_matvec_count),
        "precond_count": int(res.precond_count),
        "precond_dropped": bool(res.precond_dropped),
        "identification_steps": int(res.identification_steps),
        "requested_tol": res.requested_tol, "effective_tol": res.effective_tol,
        "validation_scope": res.validation_scope,
        "operator_error_kind": res.operator_error_kind,
        "fw_error": res.fw_error, "penetration_error": res.penetration_error,
        "load_error": res.load_error, "pk_error": res.pk_error,
        "contact_area": float(res.contact_area),
        "mean_pressure": float(res.mean_pressure),
        "active_rounds": int(res.active_rounds),
        "active_fallback": bool(res.active_fallback),
        "stage_stats": list(res.stage_stats),
        "pressure_sha": digest(res.pressure),
        "run_status": "ok",
        "timestamp": time.strftime("%Y-%m-%dT%H:%M:%S"),
    }


def append(path: str, rec: dict, prov: dict = None) -> None:
    """Append one record (with provenance) as a JSON line."""
    rec = dict(rec)
    rec["provenance"] = prov if prov is not None else provenance()
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "a") as f:
        f.write(json.dumps(rec) + "\n")
