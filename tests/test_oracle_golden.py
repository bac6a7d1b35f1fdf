"""Pin the CPU restatement (the oracle) to the compiled reference: bit-identical golden vectors."""
import numpy as np
import pytest

from golden_util import counted_cases, counted_module_indices, golden_names, load_golden, load_progress, progress_counts
from oracle.oracle import OracleDomain

CASES = golden_names()


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64)) or np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference(case):
    g = load_golden(case)
    dom = OracleDomain(g.spec())
    try:
        # frame 0: state after populateVariablesFromState (equationset.cpp:96-104)
        for name, ref in g.frame(0).items():
            assert _bits_equal(dom.grid(name), ref), "%s: setup mismatch in %s" % (case, name)
        step = 0
        for k in g.frames[1:]:
            dts = dom.step(k - step)
            np.testing.assert_array_equal(dts, g.dt_steps[step:k], err_msg="%s: dt sequence" % case)
            step = k
            for name, ref in g.frame(k).items():
                got = dom.grid(name)
                if not _bits_equal(got, ref):
                    diff = np.nanmax(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))
                    pytest.fail("%s: frame %d grid %s differs (max rel %.3e)" % (case, k, name, diff))
    finally:
        dom.close()


@pytest.mark.parametrize("case", counted_cases())
def test_oracle_subcycle_counts_match_reference_progress(case):
    """The integer subcycle counts of every step equal the ones the reference printed for it
    (fileio.cpp:283-295), module by module."""
    g = load_golden(case)
    lines = load_progress(case)
    assert len(lines) == g.nsteps
    spec = g.spec()
    idx = counted_module_indices(spec)
    dom = OracleDomain(spec)
    try:
        for s, line in enumerate(lines):
            dom.step(1)
            got = [dom.subcycles()[k] for k in idx]
            assert got == progress_counts(line), "%s step %d: %s vs reference %r" % (case, s, got, line)
    finally:
        dom.close()
