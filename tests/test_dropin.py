"""The C++ drop-in API (include/mmprop/*.hpp over libmmprop.so), driven by the programs
under tests/cpp the way the reference's simulations/main.cpp drives its solver."""
import os
import subprocess

import numpy as np
import pytest

from mm_propagation_amd import configs

from oracle_binding import Oracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "cpp", "_build")


def _binary(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (python -m mm_propagation_amd.build)")
    return path


def test_dropin_containers_and_tableaux():
    """Host-only: ComplexArray/ComplexArraysContainer arithmetic, Sparse3DArray insertion
    order and the Butcher tableaux (row sums = c_i, sum b = 1)."""
    r = subprocess.run([_binary("dropin_selftest")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


def _two_mode_params(nt):
    """The same problem tests/cpp/dropin_propagate.cpp builds with the drop-in classes."""
    return configs.two_mode_gaussian(nt=nt)


@pytest.mark.gpu
@pytest.mark.parametrize("raman", [True, False])
def test_dropin_program_matches_oracle(tmp_path, raman):
    nt, steps = 1024, 60
    out = tmp_path / "fields.bin"
    cmd = [_binary("dropin_propagate"), str(nt), str(steps), str(out)] + ([] if raman else ["raman_off"])
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=np.float64).reshape(2, nt, 2)
    got = got[..., 0] + 1j * got[..., 1]
    orc = Oracle(_two_mode_params(nt), raman=raman, pair_decode=1)  # the GPU classes decode by pair value
    orc.compute_lawson_rk(steps)
    ref = orc.get_result()
    err = configs.rel_l2_per_mode(got, ref)
    assert np.all(err <= 1e-10), err


@pytest.mark.gpu
def test_dropin_batch_and_server_equal_the_solves_alone():
    """tests/cpp/dropin_batch.cpp: MultimodePropagationBatchGPU and PropagationServer (three
    2-mode fibers, a server with two slots) against MultimodePropagationGPU_RamanON, bit for bit."""
    r = subprocess.run([_binary("dropin_batch")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
