"""The header-only C++ drop-in (include/tensor_train.hpp): it compiles
against libtt_amd.so with the host compiler (CPU), and its re-expression of
the reference's hot-path C++ tests passes on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "_build", "test_dropin")


def test_dropin_compiles():
    subprocess.run(["make", "-s", "-B", "-C", CPP], check=True)
    assert os.path.exists(BIN)


def test_host_value_types():
    # Host-only parts of the header (to_dense of tt_matrix): no device call.
    host = os.path.join(CPP, "_build", "test_host_types")
    if not os.path.exists(host):
        subprocess.run(["make", "-s", "-C", CPP], check=True)
    r = subprocess.run([host], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_dropin_runs_on_gpu():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", CPP], check=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
