"""pytest configuration: the `gpu` marker and shared fixtures.

The oracle libraries (oracle/) are test infrastructure: only tests, smoke()
and the benchmark's CPU-baseline leg load them.
"""
import os
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

# the hip backend looks every program up in its tuned-schedule store: tests
# start from an empty one (tests that tune use directories of their own)
os.environ.setdefault("LT_HIP_TUNED_DIR", tempfile.mkdtemp(prefix="lt_tuned_session_"))

ORACLE_LIB = os.path.join(ROOT, "oracle", "liblt_oracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libloop_tool_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X)")


@pytest.fixture(scope="session")
def lt():
    import loop_tool_amd as lt_mod
    if "cpp" not in lt_mod.backends():
        lt_mod.load_lib(ORACLE_LIB)
    # lazy `.numpy()` in tests resolves on the oracle; hip runs are explicit
    lt_mod.set_default_backend("cpp")
    return lt_mod.lib


@pytest.fixture(scope="session")
def ref_opt():
    """The reference library built from its sources (oracle/_ref), or None
    where it is not built (no /root/reference)."""
    if not os.path.exists(REF_LIB):
        return None
    from loop_tool_amd import capi
    # the reference's C backend compiles to fixed /tmp names: one compile at a
    # time across the test processes
    return capi.Library(REF_LIB, compile_lock="/tmp/lt_ref_cpp_backend.lock")


@pytest.fixture(scope="session")
def ref(ref_opt):
    if ref_opt is None:
        pytest.skip("reference library oracle/_ref not built")
    return ref_opt


@pytest.fixture(scope="session")
def hip_dev():
    import loop_tool_amd.hip as hip
    if not hip.available():
        pytest.fail("no HIP device visible: GPU tests need the MI355X and the built "
                    "loop_tool_amd/liblt_amd.so", pytrace=False)
    hip.init()
    return hip
