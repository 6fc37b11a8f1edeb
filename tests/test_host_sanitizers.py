"""Host-only ABI code under the address and undefined-behaviour sanitizers
(CPU only): `make asan` builds smaiso-wcomp_amd/csrc/host_abi.cpp (parameter
validation, the interface IC and its VJP, the reader of the reference's
psi<rank>.dat files, checkpoint headers) and the driver's argument parser
with g++ -fsanitize=address,undefined, together with host/abi_host_check.cpp,
which feeds them valid and malformed inputs (N = 1 grids, truncated and
non-numeric psi files, foreign and corrupt checkpoints, bad argv)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "smaiso-wcomp_amd")


def test_host_abi_under_sanitizers():
    b = subprocess.run(["make", "-C", PKG, "asan"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    exe = os.path.join(PKG, "build", "abi_host_check_asan")
    # verify_asan_link_order=0: a library preloaded by the environment may come
    # before the ASan runtime in the load order
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "host ABI checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
