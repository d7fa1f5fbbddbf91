"""The host set-up records the same step whatever its thread count: with one
thread every least-squares fit is computed when the recorder first asks for
it; with more, a first pass only collects the requests (first call wins, as
in the reference's point cache), the fits run in parallel, and a second pass
records the rows.  The digest covers every recorded host table (groups, slot
tables, pressure and Helmholtz systems), bit for bit."""
import pytest

from higflow_amd import cases, native


@pytest.mark.parametrize("make", [
    lambda: cases.wall_layer_cavity(4, dim=3, levels=2),
    lambda: cases.contraction(levels=2, h0=0.5, length=8.0, tempdiscrtype=cases.SEMI_IMPLICIT_EULER),
    lambda: cases.ve_cavity(16, tempdiscrtype=cases.SEMI_IMPLICIT_EULER),
])
def test_setup_tables_independent_of_threads(make, monkeypatch):
    monkeypatch.setenv("HFH_HOST_ONLY", "1")
    digests = []
    for nth in ("1", "2", "8"):
        monkeypatch.setenv("HFH_SETUP_THREADS", nth)
        digests.append(native.Flow(make()).tables_digest())
    assert digests[0] == digests[1] == digests[2], digests


def test_digest_sees_a_different_step(monkeypatch):
    # the digest is not constant: another time step changes the Helmholtz rows
    monkeypatch.setenv("HFH_HOST_ONLY", "1")
    a = native.Flow(cases.ve_cavity(16, dt=1e-4, tempdiscrtype=cases.SEMI_IMPLICIT_EULER)).tables_digest()
    b = native.Flow(cases.ve_cavity(16, dt=2e-4, tempdiscrtype=cases.SEMI_IMPLICIT_EULER)).tables_digest()
    assert a != b
