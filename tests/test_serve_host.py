"""Host-side behaviour of the served loop that needs no GPU: start, stats
and stop of an idle server, and argument errors."""
import ctypes as C

import pytest

from computeql_amd.native import QE_ServerStats, load


@pytest.fixture()
def lib():
    try:
        return load()
    except Exception as e:  # library not built in this checkout
        pytest.skip(str(e))


def test_idle_server_start_stats_stop(lib):
    srv = lib.qe_server_start(None)
    assert srv
    st = QE_ServerStats()
    lib.qe_server_stats(srv, C.byref(st))
    assert st.submitted == 0 and st.completed == 0 and st.p99_ns == 0
    lib.qe_server_stats_reset(srv)
    lib.qe_server_stop(srv)


def test_submit_rejects_null_query(lib):
    srv = lib.qe_server_start(None)
    assert lib.qe_server_submit(srv, None, None) == 0
    assert b"null argument" in lib.qe_last_error()
    lib.qe_server_stop(srv)
