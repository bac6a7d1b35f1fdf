"""Two-fluid (ideal_2F) GPU path through the C ABI: bit-identical to the CPU restatement on every
boundary kind and integrator, and the batched mode (one launch per kernel for many domains, HIP
graph replay, autotuned block shape) bit-identical to stepping each domain on its own."""
import numpy as np
import pytest

from oracle.oracle import OracleDomain
from spruce_amd import workloads as W
from spruce_amd.gpu import GpuBatch, GpuDomain
from spruce_amd.spec import spec_from_files

pytestmark = pytest.mark.gpu

EVOLVED = ["i_rho", "e_rho", "i_mom_x", "i_mom_y", "e_mom_x", "e_mom_y", "i_thermal_energy", "e_thermal_energy",
           "E_x", "E_y", "E_z", "bi_x", "bi_y", "bi_z"]
DERIVED = ["i_temp", "e_temp", "i_n", "e_n", "i_v_x", "e_v_y", "j_x", "j_y", "press", "rho_c", "dn", "dt", "dt_i",
           "b_mag", "b_hat_x", "curlE_z", "divE", "divB", "i_dPdx", "e_dPdx"]


def same_bits(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and (np.array_equal(a.view(np.int64), b.view(np.int64)) or np.array_equal(a, b, equal_nan=True))


def describe(a, b):
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    k = np.unravel_index(np.nanargmax(d), d.shape)
    return "max rel diff %.3e at %s (%r vs %r), %d cells differ" % (d[k], k, a[k], b[k], int(np.sum(a != b)))


def tf_spec(nx, ny, seed, bc, integ, curl, **kw):
    st = W.ucnp_state(nx, ny, seed=seed, v_noise=kw.pop("v_noise", 2e4), e_noise=kw.pop("e_noise", 1e-3),
                      n_peak=kw.pop("n_peak", 5e9))
    cfg = W.ucnp_config(steps=10, integrator=integ, remove_curl_terms=not curl, bc=bc, **kw)
    return spec_from_files(cfg, st)


CASES = [
    ("ucnp_rk2_curl", 72, 56, ("open_ucnp",) * 4, "rk2", True, {}),
    ("periodic_rk4_curl", 64, 48, ("periodic",) * 4, "rk4", True, {}),
    ("open_euler", 56, 72, ("open",) * 4, "euler", False, dict(open_boundary_strength=0.5)),
    ("reflect_rk2", 60, 52, ("reflect",) * 4, "rk2", False, {}),
    ("fixed_rk4_curl", 52, 60, ("fixed",) * 4, "rk4", True, {}),
    ("mixed_rk2", 64, 64, ("periodic", "periodic", "open", "open_ucnp"), "rk2", False,
     dict(open_boundary_strength=0.3, open_boundary_decay_base=0.7)),
]


@pytest.mark.parametrize("block", ["64x4", "strip4", "strip8"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_twofluid_matches_oracle(case, block):
    """Both stage kernels: one thread per cell (64x4 blocks) and wavefront strips of 4 and 8 rows."""
    name, nx, ny, bc, integ, curl, kw = case
    print("case", name, nx, ny, block, flush=True)
    spec = tf_spec(nx, ny, 11, bc, integ, curl, **kw)
    ref = OracleDomain(spec)
    gpu = GpuDomain(spec, options={"block": block})
    try:
        for v in EVOLVED + DERIVED:
            a, b = gpu.grid(v), ref.grid(v)
            assert same_bits(a, b), "%s setup %s: %s" % (name, v, describe(a, b))
        r = ref.rhs()
        g = gpu.rhs()
        for k, v in enumerate(EVOLVED):
            assert same_bits(g[k], r[k]), "%s rhs %s: %s" % (name, v, describe(g[k], r[k]))
        for chunk in (1, 3):
            dg = gpu.step(chunk)
            dr = ref.step(chunk)
            assert same_bits(dg, dr), "%s dt %s vs %s" % (name, dg, dr)
            for v in EVOLVED + DERIVED:
                a, b = gpu.grid(v), ref.grid(v)
                assert same_bits(a, b), "%s after %d steps %s: %s" % (name, chunk, v, describe(a, b))
    finally:
        gpu.close()
        ref.close()


SIZES = [(64, 64), (96, 80), (128, 128), (48, 112), (80, 96), (72, 72)]


def _members(integ="rk2", curl=False, sizes=SIZES, bc=("open_ucnp",) * 4):
    return [tf_spec(nx, ny, 100 + k, bc, integ, curl, n_peak=1e9 * (1 + k)) for k, (nx, ny) in enumerate(sizes)]


def _run_alone(specs, chunks):
    out, dts = [], []
    for s in specs:
        d = GpuDomain(s)
        dt = np.concatenate([d.step(c) for c in chunks])
        out.append({v: d.grid(v) for v in EVOLVED + ["dt"]})
        dts.append(dt)
        d.close()
    return out, np.array(dts)


def _check_same(doms, alone, dts_b, dts_a, what):
    assert same_bits(dts_b, dts_a), "%s: dt %s vs %s" % (what, dts_b, dts_a)
    for k, d in enumerate(doms):
        for v in EVOLVED + ["dt"]:
            a, b = d.grid(v), alone[k][v]
            assert same_bits(a, b), "%s member %d %s: %s" % (what, k, v, describe(a, b))


@pytest.mark.parametrize("integ,curl,graph", [("rk2", False, True), ("rk2", True, False), ("rk4", True, True),
                                              ("euler", False, True)])
def test_batch_matches_individual_runs(integ, curl, graph):
    print("batch", integ, curl, graph, flush=True)
    specs = _members(integ, curl)
    chunks = [3, 1, 4]
    alone, dts_a = _run_alone(specs, chunks)
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms, graph=graph)
    try:
        dts_b = []
        for c in chunks:
            dts_b.append(b.step(c))
            # eager work between graph replays
            for d in doms:
                d.grid("e_rho")
        _check_same(doms, alone, np.concatenate(dts_b, axis=1), dts_a, "batch %s" % integ)
    finally:
        b.close()
        for d in doms:
            d.close()


def test_batch_mixed_boundaries_and_blocks():
    print("batch mixed", flush=True)
    specs = [tf_spec(64, 80, 7, ("periodic", "periodic", "open", "open_ucnp"), "rk2", False),
             tf_spec(56, 56, 8, ("reflect",) * 4, "rk2", False),
             tf_spec(72, 48, 9, ("fixed",) * 4, "rk2", True),
             tf_spec(40, 96, 10, ("open",) * 4, "rk2", False, open_boundary_strength=0.4)]
    alone, dts_a = _run_alone(specs, [5])
    for block in ("16x4", "32x8", "128x2", "strip4", "strip8", "strip16"):
        doms = [GpuDomain(s) for s in specs]
        b = GpuBatch(doms, graph=True, block=block)
        try:
            dts_b = b.step(5)
            _check_same(doms, alone, dts_b, dts_a, "block %s" % block)
        finally:
            b.close()
            for d in doms:
                d.close()


def test_autotune_keeps_state_and_results():
    print("autotune", flush=True)
    specs = _members("rk2", False, sizes=[(128, 128), (64, 96), (192, 160)])
    alone, dts_a = _run_alone(specs, [4])
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms, graph=True)
    try:
        before = [{v: d.grid(v) for v in EVOLVED} for d in doms]
        t0 = [d.time for d in doms]
        shape, ms = b.autotune(2)
        print("autotune chose", shape, ms, flush=True)
        assert ms > 0 and "x" in shape
        for k, d in enumerate(doms):
            assert d.time == t0[k]
            for v in EVOLVED:
                assert same_bits(d.grid(v), before[k][v]), "autotune changed %s of member %d" % (v, k)
        dts_b = b.step(4)
        _check_same(doms, alone, dts_b, dts_a, "after autotune")
    finally:
        b.close()
        for d in doms:
            d.close()


def test_domain_in_batch_cannot_step_alone_and_recovers():
    specs = _members("rk2", False, sizes=[(64, 64), (48, 80)])
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms)
    try:
        with pytest.raises(Exception):
            doms[0].step(1)
        b.step(2)
    finally:
        b.close()
    # after the batch is gone the domain steps alone again, continuing from the batch's state
    d2 = GpuDomain(specs[0])
    d2.step(2)
    dt_a = d2.step(1)
    dt_b = doms[0].step(1)
    assert same_bits(dt_b, dt_a)
    for v in EVOLVED:
        assert same_bits(doms[0].grid(v), d2.grid(v))
    d2.close()
    for d in doms:
        d.close()


def test_autotune_with_more_trial_steps_than_the_history_holds():
    """Fresh domains hold two step sizes; autotuning times 12 steps per launch shape.  The
    library grows every member's step-size history first and reports an overflow as an error, so
    this passes only if no step size was dropped or written past the history."""
    print("autotune 12 trial steps", flush=True)
    specs = _members("rk2", False, sizes=[(64, 64), (48, 80)])
    alone, dts_a = _run_alone(specs, [3])
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms, graph=True)
    try:
        b.autotune(12)
        dts_b = b.step(3)
        _check_same(doms, alone, dts_b, dts_a, "after autotune(12)")
    finally:
        b.close()
        for d in doms:
            d.close()


def test_batch_of_64_expansions_matches_individual_runs():
    """The batched-mode workload itself: 64 two-fluid expansions with edges from 128 to 512, one
    batch with graph replay and the autotuned launch shape, 3 steps against each domain stepped
    alone."""
    print("batch of 64", flush=True)
    members = W.batch_2f_members(64)
    specs = [spec_from_files(c, s) for s, c in members]
    alone, dts_a = _run_alone(specs, [3])
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms, graph=True)
    try:
        shape, _ = b.autotune(2)
        print("batch of 64: shape", shape, flush=True)
        dts_b = b.step(3)
        _check_same(doms, alone, dts_b, dts_a, "64-domain batch")
    finally:
        b.close()
        for d in doms:
            d.close()


@pytest.mark.parametrize("integ,curl", [("rk2", False), ("rk4", True), ("euler", False)])
def test_persistent_batch_matches_individual_runs(integ, curl):
    """Every step of a call in one cooperative launch (grid barriers between the phases): bit for
    bit the members stepped alone, with eager work between calls."""
    print("persistent batch", integ, curl, flush=True)
    specs = _members(integ, curl)
    chunks = [3, 1, 4]
    alone, dts_a = _run_alone(specs, chunks)
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms, graph=False)
    try:
        b.set_option("persistent", "true")
        dts_b = []
        for c in chunks:
            dts_b.append(b.step(c))
            for d in doms:
                d.grid("e_rho")
        _check_same(doms, alone, np.concatenate(dts_b, axis=1), dts_a, "persistent batch %s" % integ)
    finally:
        b.close()
        for d in doms:
            d.close()


def test_persistent_batch_mixed_boundaries_and_shapes():
    """Every boundary kind beside each other; a strip shape is kept, a cell shape gives 4-row strips."""
    print("persistent batch mixed", flush=True)
    specs = [tf_spec(64, 80, 7, ("periodic", "periodic", "open", "open_ucnp"), "rk2", False),
             tf_spec(56, 56, 8, ("reflect",) * 4, "rk2", False),
             tf_spec(72, 48, 9, ("fixed",) * 4, "rk2", True),
             tf_spec(40, 96, 10, ("open",) * 4, "rk2", False, open_boundary_strength=0.4)]
    alone, dts_a = _run_alone(specs, [5])
    for block in ("strip8", "32x8"):
        doms = [GpuDomain(s) for s in specs]
        b = GpuBatch(doms, graph=True, block=block)
        try:
            b.set_option("persistent", "true")
            dts_b = b.step(5)
            _check_same(doms, alone, dts_b, dts_a, "persistent block %s" % block)
        finally:
            b.close()
            for d in doms:
                d.close()


def test_persistent_batch_of_64_expansions_matches_individual_runs():
    print("persistent batch of 64", flush=True)
    members = W.batch_2f_members(64)
    specs = [spec_from_files(c, s) for s, c in members]
    alone, dts_a = _run_alone(specs, [3])
    doms = [GpuDomain(s) for s in specs]
    b = GpuBatch(doms, graph=True, block="strip4")
    try:
        b.set_option("persistent", "true")
        dts_b = b.step(3)
        _check_same(doms, alone, dts_b, dts_a, "persistent 64-domain batch")
    finally:
        b.close()
        for d in doms:
            d.close()
