"""GPU timestep against the reference golden vectors and the CPU oracle.

The host driver (pawsim_amd/bin/pawsim_gpu, all arithmetic in HIP kernels
behind the C ABI) runs each golden input; with the bit-replica solvers (SOR,
gathered multigrid) U/V/H/P must equal the reference's outputs bit for bit
and every solve must take the same number of iterations.
"""
import os

import numpy as np
import pytest

import pawsim_amd
from conftest import golden_ids, load_golden, parse_iters, run_exe, same_bits
from pawsim_amd import inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_ids())
def test_gpu_matches_reference_golden(name, tmp_path):
    z, setup = load_golden(name)
    infile = setup.write(str(tmp_path / "in"))
    out = str(tmp_path / "gpu")
    print(f"[gpu parity] {name} {setup.nx}x{setup.ny}x{setup.nlay}", flush=True)
    stdout = run_exe(pawsim_amd.DRIVER_PATH, infile, out, timeout=300)
    for n in z["saves"]:
        got = inputs.read_outputs(out, setup.nx, setup.ny, setup.nlay, int(n))
        for key, arr in got.items():
            ref = z[f"{key}_n{n}"]
            assert same_bits(arr, ref), f"{key} at save {n}: max|diff|={np.max(np.abs(arr - ref)):.3e}"
    assert parse_iters(stdout) == [int(x) for x in z["solves"][:, 1]]
