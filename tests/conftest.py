import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "htr-solver_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI library)")
    config.addinivalue_line("markers", "slow: longer CPU test")
    config._htr_parity = []


@pytest.fixture
def parity_record(request):
    """Collects one line per Euler parity check (faces, TENO faces, limited
    faces, decision flips per kind, errors) for the end-of-session table."""
    rows = request.config._htr_parity

    def record(label, d, nfaces, teno, limited, kinds, err, err_all):
        rows.append((label, d, nfaces, teno, limited) + tuple(kinds) + (err, err_all))
    return record


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    rows = getattr(config, "_htr_parity", [])
    if not rows:
        return
    tr = terminalreporter
    tr.section("Euler decision flips (GPU vs oracle)")
    tr.write_line("%-26s %3s %7s %6s %7s %6s %6s %6s %10s %10s" % (
        "case", "dir", "faces", "TENO", "limited", "hyb", "lim", "teno", "err", "err_all"))
    for r in rows:
        tr.write_line("%-26s %3d %7d %6d %7d %6d %6d %6d %10.2e %10.2e" % r)
