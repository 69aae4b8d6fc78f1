"""GPU: the batched workload in resumable steps (rf_queue_*). A queue run in
steps, saved after the first, loaded into a model with another lane count
and finished there gives the hashes of the one-call batched form and of the
oracle, in the per-lane persistent and bit-plane forms, and with per-stimulus
memory images (RV32I)."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry, rv32i
from rtlflow_amd.model import MODE_BITPLANE, MODE_FUSED, RFError

SEED = registry.SEED_CFG[2]


def _budgets(n, seed):
    rng = np.random.default_rng(seed)
    b = rng.integers(0, 90, size=n).astype(np.uint32)
    b[::37] = 0
    return b


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_FUSED, MODE_BITPLANE])
def test_queue_resumes_on_another_model(gpu, mode):
    name, nstim = "net_small", 600
    cd = registry.get(name)
    budgets = _budgets(nstim, 11 + mode)
    ref = oracle.netlist_run(registry.table_bytes(name), SEED, nstim=nstim, budgets=budgets)
    with cd.model(128) as m:
        whole = m.submit_queue(SEED, budgets, mode=mode)
    order = np.argsort(-budgets.astype(np.int64), kind="stable")   # longest budget first
    with cd.model(128) as m:
        m.queue_open(SEED, budgets)
        assert m.queue_step(250, mode) == 250
        part = m.queue_hashes()
        blob = m.queue_save()
    ran = np.zeros(nstim, bool)
    ran[order[:250]] = True
    assert (part[ran] == ref[ran]).all()
    assert (part[~ran] == 0).all()
    with cd.model(192) as m2:
        m2.queue_load(blob)
        assert m2.queue_step(100, mode) == 350
        assert m2.queue_step(10 ** 6, mode) == nstim
        assert m2.queue_step(5, mode) == nstim
        resumed = m2.queue_hashes()
    assert (whole == ref).all()
    assert (resumed == ref).all()


@pytest.mark.gpu
def test_queue_resume_with_memory_images(gpu):
    nstim = 300
    seed = registry.SEED_CFG[3]
    progs = rv32i.programs(seed, np.arange(nstim))
    budgets = rv32i.budgets(seed, nstim, median=400, lo=50, hi=1500)
    ref = oracle.rv32i_run(progs, seed, budgets)
    cd = registry.get("rv32i")
    with cd.model(64) as m:
        m.set_mem_images("imem", progs)
        m.queue_open(seed, budgets)
        assert m.queue_step(120) == 120
        blob = m.queue_save()
    with cd.model(96) as m2:
        m2.set_mem_images("imem", progs)
        m2.queue_load(blob)
        assert m2.queue_step(nstim) == nstim
        got = m2.queue_hashes()
    assert (got == ref).all(), f"{int((got != ref).sum())} of {nstim} differ"


@pytest.mark.gpu
def test_queue_errors(gpu):
    cd = registry.get("net_small")
    with cd.model(64) as m:
        with pytest.raises(RFError, match="no open queue"):
            m.queue_step(10)
        m.queue_open(SEED, _budgets(40, 1))
        m.queue_step(10)
        blob = m.queue_save()
        with pytest.raises(RFError, match="RF_MODE_FUSED or RF_MODE_BITPLANE"):
            m.queue_step(10, 0)
    with registry.get("net_tiny").model(64) as m2:
        with pytest.raises(RFError, match="another design"):
            m2.queue_load(blob)
        with pytest.raises(RFError, match="magic"):
            m2.queue_load(b"\0" * len(blob))
